// C++ names for the library's plain data types.
//
// They replace CUDA's float3/float4/int3 (vector_types.h) and the reference
// PODs SimulationParams (include/pbf/simulation_params.hpp:6-29) and the
// colliders (include/pbf/colliders.hpp:7-25).  Layouts are byte-identical,
// so a reference caller only switches includes and namespace.
#pragma once

#include <stdexcept>

#include "pbf_amd.h"

namespace pbf {

using float3 = pbf_float3;
using float4 = pbf_float4;
using int3 = pbf_int3;

using SimulationParams = pbf_sim_params;
using Container = pbf_container;
using SphereCollider = pbf_sphere;
using BoxCollider = pbf_box;
using PlaneCollider = pbf_plane;

inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int3 make_int3(int x, int y, int z) { return int3{x, y, z}; }

static_assert(sizeof(float3) == 12 && alignof(float3) == 4, "float3 layout");
static_assert(sizeof(float4) == 16 && alignof(float4) == 16, "float4 layout");
static_assert(sizeof(SimulationParams) == 72, "SimulationParams layout");
static_assert(sizeof(Container) == 24 && sizeof(SphereCollider) == 16, "collider layout");
static_assert(sizeof(BoxCollider) == 24 && sizeof(PlaneCollider) == 24, "collider layout");

// Thrown where the reference throws std::runtime_error("Collision constraints
// did not converge") (src/pbf/collision_system.cu:171); still a runtime_error.
class collision_nonconvergence : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Converts a status of the C ABI into the exception the reference throws.
void throw_status(pbf_status status);

inline void check(pbf_status status) {
    if (status != PBF_OK) throw_status(status);
}

}  // namespace pbf
