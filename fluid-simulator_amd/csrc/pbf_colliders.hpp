// Device view of a collider set: device pointers plus counts.  Plain data,
// shared by host code and kernels.
#pragma once

#include "pbf_amd.h"

namespace pbfdev {

struct Colliders {
    const pbf_container* container;  // device pointer, never null inside the solver
    const pbf_sphere* spheres;
    const pbf_box* boxes;
    const pbf_plane* planes;
    int sphereCount;
    int boxCount;
    int planeCount;
};

}  // namespace pbfdev
