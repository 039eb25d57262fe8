// Container, sphere, box and plane collisions.
//
// Replaces CollisionSystem (include/pbf/collision_system.hpp:11-57,
// src/pbf/collision_system.cu:29-277): position projection to a fixed point
// in at most 32 passes, velocity contact response against the incoming
// velocity, the same validation and strong guarantees.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "pbf/types.hpp"

namespace pbf {

class CollisionSystem {
public:
    CollisionSystem();
    explicit CollisionSystem(const Container& container);
    ~CollisionSystem();
    CollisionSystem(CollisionSystem&&) noexcept;
    CollisionSystem& operator=(CollisionSystem&&) noexcept;
    CollisionSystem(const CollisionSystem&) = delete;
    CollisionSystem& operator=(const CollisionSystem&) = delete;

    void setContainer(const Container& container);

    void setSpheres(const std::vector<SphereCollider>& spheres);
    void clearSpheres();
    void setBoxes(const std::vector<BoxCollider>& boxes);
    void clearBoxes();
    void setPlanes(const std::vector<PlaneCollider>& planes);
    void clearPlanes();

    // Projects device positions; throws collision_nonconvergence (a
    // std::runtime_error) when 32 passes still correct something.
    void solve(float4* predictedPositions, std::size_t particleCount, float particleRadius);

    void resolveVelocities(const float4* positions, float4* velocities, std::size_t particleCount,
                           float particleRadius, float restitution, float friction);
    void resolveVelocities(const float4* positions, const float4* incomingVelocities,
                           float4* velocities, std::size_t particleCount, float particleRadius,
                           float restitution, float friction);

    struct State;  // library-internal
    State& state() { return *state_; }
    const State& state() const { return *state_; }

private:
    std::unique_ptr<State> state_;
};

}  // namespace pbf
