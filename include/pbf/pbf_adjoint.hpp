// Reverse-mode derivatives of PBFSolver frames.  No counterpart in the
// reference (its solver has no derivatives).
//
// forward() runs frames of a solver and keeps its working set every
// `checkpointEvery` frames, and records a tape (neighbour lists, orderings
// and per-pass records of every substep) of the last segment as it runs it;
// backward() runs the hand-written adjoint kernels back through that tape,
// then replays each earlier segment from its checkpoint with the tape on and
// goes back through it, so memory holds one segment's tape at a time.  With
// checkpointEvery >= frameCount nothing is replayed.  Neighbour lists and every
// discrete decision of the forward (which collision projections act and
// their branches, which contacts act) are held fixed; between them the
// derivatives are exact up to FP32 rounding.  The collision solve and the
// contact velocities take the container, spheres, rounded boxes and planes.
// XSPH viscosity and vorticity strength may be zero (the passes are then
// skipped, as in the reference, and their derivatives are the one-sided ones
// at zero); under the adjoint every substep runs the post-solve grid build,
// which the reference skips when both are zero (the results differ from a
// plain run only in the summation order of the next substep).
#pragma once

#include <cstddef>
#include <memory>

#include "pbf/pbf_solver.hpp"
#include "pbf/types.hpp"

namespace pbf {

class PBFAdjoint {
public:
    // The solver must outlive this object and keep its particles and
    // parameters between forward() and backward().
    explicit PBFAdjoint(PBFSolver& solver);
    ~PBFAdjoint();
    PBFAdjoint(const PBFAdjoint&) = delete;
    PBFAdjoint& operator=(const PBFAdjoint&) = delete;

    // Runs frameCount frames (as PBFSolver::run) from the solver's current
    // state, keeping a checkpoint every checkpointEvery frames.
    void forward(std::size_t frameCount, std::size_t checkpointEvery);

    // Given dL/dx and dL/dv of the state forward() ended in (host arrays in
    // API order, particleCount records; null means zero), writes dL/dx and
    // dL/dv of the state it started from, and dL/d(xsphViscosity) and
    // dL/d(vorticityStrength) (any output may be null).  The solver is left
    // in the state forward() ended in.
    void backward(const float4* xBar, const float4* vBar, std::size_t particleCount,
                  float4* x0Bar, float4* v0Bar, double* viscosityBar, double* vorticityBar);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace pbf
