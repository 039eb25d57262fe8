// Many independent PBF scenes advanced together: one launch per stage of the
// step for all scenes.  No counterpart in the reference, whose PBFSolver
// holds one scene (include/pbf/pbf_solver.hpp:15-33); a reference caller
// with many scenes loops PBFSolver::step() over them.
//
// Every scene keeps its own grid (its bounds, cell size = smoothing radius)
// and its own container (its bounds, open at the top); the simulation
// parameters are shared.  The scenes share one working set: cell keys are
// offset per scene, so one sort puts every scene in cell order inside its
// own slot range, and the list build, constraint passes and contacts look
// up the scene of a slot in device tables.  The results of each scene are
// bitwise those of a PBFSolver that holds that scene alone.
#pragma once

#include <cstddef>
#include <vector>

#include "pbf/pbf_solver.hpp"
#include "pbf/types.hpp"

namespace pbf {

struct BatchScene {
    std::size_t particleCount;
    float3 minBounds;
    float3 maxBounds;
};

class PBFBatchSolver {
public:
    PBFBatchSolver();
    ~PBFBatchSolver();
    PBFBatchSolver(PBFBatchSolver&&) noexcept;
    PBFBatchSolver& operator=(PBFBatchSolver&&) noexcept;
    PBFBatchSolver(const PBFBatchSolver&) = delete;
    PBFBatchSolver& operator=(const PBFBatchSolver&) = delete;

    // Validates like PBFSolver::initialize per scene (the parameters once,
    // the bounds and capacity of every scene) and allocates for all of them.
    void initialize(const std::vector<BatchScene>& scenes, const SimulationParams& params);
    // All particles, scene after scene (particleCount = the sum of the
    // scenes' counts); validated like PBFSolver::setParticles.
    void setParticles(const float4* positions, const float4* velocities, std::size_t particleCount);
    void step();
    void run(std::size_t frameCount);
    // All particles in the order of setParticles.
    void copyPositionsToHost(float4* out, std::size_t particleCount) const;
    void copyVelocitiesToHost(float4* out, std::size_t particleCount) const;

    std::size_t sceneCount() const;
    std::size_t particleCount() const;
    // First particle of a scene in the order of setParticles.
    std::size_t sceneFirst(std::size_t scene) const;

    void setGraphs(bool enabled);
    void setIncrementalSort(bool enabled);  // PBFSolver::setIncrementalSort
    void setFusedXsph(bool enabled);        // PBFSolver::setFusedXsph
    // Skin lists for every scene (PBFSolver::setNeighborSkin).  A frame in
    // which any scene's lists go stale runs again for every scene with the
    // reference's builds.
    void setNeighborSkin(float skin);
    std::size_t staleFrames() const;
    // Launch tunings and the autotuner of PBFSolver, over the whole batch.
    void setTuning(const PBFSolver::Tuning& tuning);
    PBFSolver::Tuning tuning() const;
    std::vector<PBFSolver::TuningTime> autotune(
        std::size_t frames, const std::vector<PBFSolver::Tuning>& candidates = {});
    void setProfiling(bool enabled);
    std::vector<PBFSolver::KernelTime> kernelTimes() const;

private:
    PBFSolver solver_;
};

}  // namespace pbf
