// Position Based Fluids solver facade.
//
// Replaces PBFSolver (include/pbf/pbf_solver.hpp:13-70,
// src/pbf/pbf_solver.cu:23-379): same member functions, argument meaning and
// exception types.  step() runs on the device without the reference's
// per-phase host synchronisations; neighbour overflow and collision
// non-convergence are collected in a device status word and thrown when the
// step returns (std::overflow_error / collision_nonconvergence).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "pbf/types.hpp"

namespace pbf {

class PBFSolver {
public:
    PBFSolver();
    ~PBFSolver();
    PBFSolver(PBFSolver&&) noexcept;
    PBFSolver& operator=(PBFSolver&&) noexcept;
    PBFSolver(const PBFSolver&) = delete;
    PBFSolver& operator=(const PBFSolver&) = delete;

    void initialize(std::size_t maxParticles, float3 minBounds, float3 maxBounds,
                    const SimulationParams& params);

    void setParticles(const float4* positions, const float4* velocities,
                      std::size_t particleCount);
    // Device-resident variant; the caller guarantees finite values.
    void setParticlesDevice(const float4* positions, const float4* velocities,
                            std::size_t particleCount);

    void setSpheres(const std::vector<SphereCollider>& spheres);
    void setBoxes(const std::vector<BoxCollider>& boxes);
    void setPlanes(const std::vector<PlaneCollider>& planes);

    void step();
    void run();
    void run(std::size_t frameCount);

    void copyPositionsToHost(float4* positions, std::size_t particleCount) const;
    void copyVelocitiesToHost(float4* velocities, std::size_t particleCount) const;
    void copyPositionsToDevice(float4* positions, std::size_t particleCount) const;
    void copyVelocitiesToDevice(float4* velocities, std::size_t particleCount) const;

    std::size_t particleCount() const;
    std::size_t maxParticles() const;
    void* stream() const;  // hipStream_t

    // Per-kernel event timing of the step (see pbf_solver_set_profiling).
    void setProfiling(bool enabled);
    // Replay frames from captured HIP graphs (default off; see
    // pbf_solver_set_graphs).  No counterpart in the reference.
    void setGraphs(bool enabled);
    struct KernelTime {
        std::string name;
        double milliseconds;
        std::uint64_t launches;
    };
    std::vector<KernelTime> kernelTimes() const;

    // The last grid build's neighbour lists, slot-major with stride
    // particleCount, and the working-set positions they were built on
    // (see pbf_solver_neighbor_lists).
    void copyNeighborLists(float4* positions, int* counts, std::uint32_t* neighbors,
                           std::size_t particleCount, int maxNeighbors) const;

    struct State;  // library-internal

private:
    friend class PBFBatchSolver;  // runs the same step over a batch of scenes
    friend class PBFAdjoint;      // replays frames with a tape
    std::unique_ptr<State> state_;
};

}  // namespace pbf
