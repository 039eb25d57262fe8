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
    // Grid builds re-sort from the previous build when the particles are in
    // its order (default on; see pbf_solver_set_incremental_sort).  Results
    // are bitwise the same either way.  No counterpart in the reference.
    void setIncrementalSort(bool enabled);
    // XSPH fused into the post-solve grid build (default on; see
    // pbf_solver_set_fused_xsph).  Results are bitwise the same either way.
    // No counterpart in the reference.
    void setFusedXsph(bool enabled);
    // Skin lists: one grid build per substep instead of one per constraint
    // iteration plus the post-solve rebuild (see
    // pbf_solver_set_neighbor_skin).  skin is a fraction of the smoothing
    // radius in [0, 1]; 0 runs the reference's builds.  No counterpart in
    // the reference.
    void setNeighborSkin(float skin);
    float neighborSkin() const;
    // Frames run again with the reference's builds because their skin
    // lists went stale, since initialize.
    std::size_t staleFrames() const;

    // Launch tunings, no counterpart in the reference: the skin (the grid's
    // cell-size multiplier is 1 + neighborSkin) and the workgroup size of
    // the neighbour passes (64, 128, 192 or 256 threads; results do not
    // depend on it).
    struct Tuning {
        float neighborSkin = 0.0f;
        int passBlock = 256;
        int particlesPerThread = 1;  // of the neighbour passes: 1, 2 or 4
    };
    struct TuningTime {
        Tuning tuning;
        double msPerFrame;
    };
    void setTuning(const Tuning& tuning);
    Tuning tuning() const;
    // The autotuner: runs `frames` frames of every candidate (the default
    // set when empty: skins 0, 0.1, 0.2, 0.3 by workgroups of 64, 128,
    // 256 by 1 or 2 particles per thread) from the current state, restoring it after each, times them on
    // the host clock (stale frames run again included), and keeps the
    // fastest.  The state, and staleFrames(), are left as they were.
    std::vector<TuningTime> autotune(std::size_t frames,
                                     const std::vector<Tuning>& candidates = {});
    static std::vector<Tuning> defaultTunings();
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
