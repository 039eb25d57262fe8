// CollisionSystem: collider configuration, position projection and velocity
// contact response.
//
// Reference: src/pbf/collision_system.cu:1-277.
#include <cmath>
#include <limits>
#include <stdexcept>

#include "host_internal.hpp"
#include "pbf_internal.hpp"

namespace pbf {
namespace {

constexpr std::size_t kBlock = 256;

// collisionBlockCount (collision_system.cu:7-13).
void check_block_count(std::size_t n) {
    const std::size_t blocks = n / kBlock + (n % kBlock != 0);
    if (blocks > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::length_error("Particle count exceeds launch dimensions");
}

// validateParticleFitsContainer (collision_system.cu:16-25): width and depth
// must hold one particle diameter; the height is not checked.
void check_fits(const Container& c, float radius) {
    const double diameter = 2.0 * static_cast<double>(radius);
    const double width = static_cast<double>(c.max.x) - static_cast<double>(c.min.x);
    const double depth = static_cast<double>(c.max.z) - static_cast<double>(c.min.z);
    if (width < diameter || depth < diameter)
        throw std::invalid_argument("Particle diameter exceeds closed container dimensions");
}

bool finite3(float3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

}  // namespace

CollisionSystem::CollisionSystem() : state_(new State()) {}
CollisionSystem::CollisionSystem(const Container& container) : state_(new State()) {
    setContainer(container);
}
CollisionSystem::~CollisionSystem() = default;
CollisionSystem::CollisionSystem(CollisionSystem&&) noexcept = default;
CollisionSystem& CollisionSystem::operator=(CollisionSystem&&) noexcept = default;

void CollisionSystem::setContainer(const Container& container) {
    if (!finite3(container.min) || !finite3(container.max) || container.max.x <= container.min.x ||
        container.max.y <= container.min.y || container.max.z <= container.min.z)
        throw std::invalid_argument("Container bounds must be finite and ordered");
    State& s = *state_;
    if (s.container.size() == 0) s.container.allocate(1);
    if (s.status.size() == 0) s.status.allocate(1);
    s.container.copyFromHostToDevice(&container, 1);
    s.hostContainer = container;
    s.hasContainer = true;
}

void CollisionSystem::setSpheres(const std::vector<SphereCollider>& spheres) {
    if (spheres.empty()) {
        clearSpheres();
        return;
    }
    for (const SphereCollider& sp : spheres) {
        if (!finite3(sp.center) || !std::isfinite(sp.radius) || sp.radius <= 0.0f)
            throw std::invalid_argument("Sphere radius must be positive");
    }
    DeviceBuffer<SphereCollider> next(spheres.size());
    next.copyFromHostToDevice(spheres.data(), spheres.size());
    state_->spheres = std::move(next);
    state_->sphereCount = spheres.size();
}

void CollisionSystem::clearSpheres() {
    state_->sphereCount = 0;
    state_->spheres.allocate(0);
}

void CollisionSystem::setBoxes(const std::vector<BoxCollider>& boxes) {
    if (boxes.empty()) {
        clearBoxes();
        return;
    }
    for (const BoxCollider& b : boxes) {
        if (!finite3(b.center) || !std::isfinite(b.halfExtents.x) || b.halfExtents.x <= 0.0f ||
            !std::isfinite(b.halfExtents.y) || b.halfExtents.y <= 0.0f ||
            !std::isfinite(b.halfExtents.z) || b.halfExtents.z <= 0.0f)
            throw std::invalid_argument("Box half extents must be positive");
    }
    DeviceBuffer<BoxCollider> next(boxes.size());
    next.copyFromHostToDevice(boxes.data(), boxes.size());
    state_->boxes = std::move(next);
    state_->boxCount = boxes.size();
}

void CollisionSystem::clearBoxes() {
    state_->boxCount = 0;
    state_->boxes.allocate(0);
}

void CollisionSystem::setPlanes(const std::vector<PlaneCollider>& planes) {
    if (planes.empty()) {
        clearPlanes();
        return;
    }
    for (const PlaneCollider& p : planes) {
        const float l2 = p.normal.x * p.normal.x + p.normal.y * p.normal.y + p.normal.z * p.normal.z;
        if (!finite3(p.point) || !std::isfinite(l2) || l2 == 0.0f)
            throw std::invalid_argument("Plane normal must not be zero");
    }
    DeviceBuffer<PlaneCollider> next(planes.size());
    next.copyFromHostToDevice(planes.data(), planes.size());
    state_->planes = std::move(next);
    state_->planeCount = planes.size();
}

void CollisionSystem::clearPlanes() {
    state_->planeCount = 0;
    state_->planes.allocate(0);
}

void CollisionSystem::solve(float4* predicted, std::size_t n, float radius) {
    if (predicted == nullptr) throw std::invalid_argument("predictedPositions must not be null");
    if (n == 0) return;
    State& s = *state_;
    if (!s.hasContainer)
        throw std::logic_error("CollisionSystem container has not been configured");
    if (!std::isfinite(radius) || radius <= 0.0f)
        throw std::invalid_argument("particleRadius must be positive");
    check_fits(s.hostContainer, radius);
    check_block_count(n);
    // One launch iterates every particle to its fixed point (at most 32
    // passes); the reference loops whole passes with a host flag read.
    hipStream_t stream = nullptr;
    launch_check(pbfint::launch_zero_u32(s.status.data(), 1, stream));
    launch_check(pbfint::launch_collide_fixed_point(predicted, n, s.view(), radius,
                                                    s.status.data(), stream));
    std::uint32_t status = 0;
    hip_check(hipMemcpy(&status, s.status.data(), sizeof(status), hipMemcpyDeviceToHost));
    if (status & PBF_STATUS_COLLISION_NONCONVERGENCE)
        throw collision_nonconvergence("Collision constraints did not converge");
}

void CollisionSystem::resolveVelocities(const float4* positions, float4* velocities,
                                        std::size_t n, float radius, float restitution,
                                        float friction) {
    resolveVelocities(positions, velocities, velocities, n, radius, restitution, friction);
}

void CollisionSystem::resolveVelocities(const float4* positions, const float4* incoming,
                                        float4* velocities, std::size_t n, float radius,
                                        float restitution, float friction) {
    if (positions == nullptr || incoming == nullptr || velocities == nullptr)
        throw std::invalid_argument("positions and velocities must not be null");
    if (n == 0) return;
    const State& s = *state_;
    if (!s.hasContainer)
        throw std::logic_error("CollisionSystem container has not been configured");
    if (!std::isfinite(radius) || radius <= 0.0f)
        throw std::invalid_argument("particleRadius must be positive");
    check_fits(s.hostContainer, radius);
    if (!std::isfinite(restitution) || restitution < 0.0f || restitution > 1.0f)
        throw std::invalid_argument("restitution must be between zero and one");
    if (!std::isfinite(friction) || friction < 0.0f || friction > 1.0f)
        throw std::invalid_argument("friction must be between zero and one");
    check_block_count(n);
    launch_check(pbf_k_resolve_velocities(positions, incoming, velocities, n, s.container.data(),
                                          s.spheres.data(), s.sphereCount, s.boxes.data(),
                                          s.boxCount, s.planes.data(), s.planeCount, radius,
                                          restitution, friction, nullptr));
}

}  // namespace pbf
