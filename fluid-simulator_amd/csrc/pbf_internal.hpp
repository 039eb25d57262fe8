// Library-internal launchers of the fused step kernels and the grid build.
// The public kernel launchers are declared in pbf_amd.h.
#pragma once

#include <cstddef>
#include <cstdint>

#include "pbf_amd.h"
#include "pbf_colliders.hpp"

namespace pbfint {

// Scenes of a batched step.  Scene s keeps its own grid (bounds, dims) and
// container; its particles occupy the slots [first, end) in every order the
// step produces (cell order sorts by global key = cellOff + the scene's own
// key, so scenes stay in their slot ranges).
struct SceneGeom {
    pbf_float3 minBounds;
    pbf_int3 dims;
    uint32_t cellOff;
    uint32_t first;
    uint32_t end;
};
// Device tables of a batched step; all null for a single scene.
struct SceneTable {
    const SceneGeom* geom = nullptr;
    const pbf_container* containers = nullptr;
    const uint32_t* sceneOf = nullptr;  // scene of every slot
};

// Skin lists: one grid build per substep, around the container points of
// the predicted positions, with radius h * (1 + skin); every pass tests each
// entry against h itself.  The lists stay complete while no particle has
// moved more than half the skin from its point (the Delta-p pass checks and
// raises PBF_STATUS_LIST_STALE); the counts within h replace the counts of
// the builds they stand in for.
// A neighbour pass's launch: the skin-list bookkeeping above, the
// workgroup size (threads, a multiple of 64 up to 256) and the particles
// per thread (results depend on neither).
struct ListPass {
    const pbf_float4* p0 = nullptr;  // predicted positions the lists were built around
    float stale2 = 0.0f;             // squared displacement from their points that is too far
    int exactMax = 0;                // > 0: neighbours within h above this raise the overflow bit
    int block = 256;
    int ppt = 1;  // particles per thread (a tuning: results do not depend on it)
    // Delta-p: with keyOut set, also the cell keys of its output positions
    // in this grid (the next grid build's keys; a batched step offsets them
    // by the scene table).
    uint32_t* keyOut = nullptr;
    pbf_float3 keyMin{};
    float keyCell = 0.0f;
    pbf_int3 keyDims{};
    // Delta-p: with clearCells set, -1 into the next build's cellStart
    // (clearCount cells).
    int* clearCells = nullptr;
    size_t clearCount = 0;
    // Delta-p: with xperm set, the record x of slot i is at xperm[i] (the
    // grid builds since the last full permutation moved p* only).
    const uint32_t* xperm = nullptr;
    // Delta-p: with prePos set, also the positions after Delta-p and before
    // the collision solve (the adjoint's tape).
    pbf_float4* prePos = nullptr;
    // Lambda, XSPH and vorticity: with tapeOut set, also their output record
    // of every particle ((p*, lambda), the velocity, omega) for the
    // adjoint's tape.
    pbf_float4* tapeOut = nullptr;
};

// Grid (k_grid.hip)
long long stencil_radius(float h, float cellSize);
pbf_status launch_find_neighbors(const pbf_float4* p, const int* cs, const int* ce, pbf_int3 g,
                                 pbf_float3 mn, float cellSize, size_t n, float h,
                                 uint32_t* nbr, int* cnt, int maxN, size_t stride,
                                 uint32_t* flag, uint32_t bit, const uint32_t* guard,
                                 void* stream);
// -1 into cellStart and, unless ce is null, cellEnd.
pbf_status launch_fill_cells(int* cs, int* ce, size_t cells, void* stream);
// Cell ranges of the sorted keys.  With splitters not null, also the keys
// at multiples of the re-sort tile (resort_splitter_count(n) entries) for
// the next resort_pairs.
pbf_status launch_cell_ranges(const uint32_t* sk, int* cs, int* ce, size_t n, void* stream,
                              uint32_t* splitters = nullptr);
// The working-set permutation of a grid build: slot i takes the records of
// slot idx[i].  With x null only p* moves, and permOut[i] = permIn[idx[i]]
// (idx[i] without permIn) records where the other records of slot i still
// are.  With x set every record moves; x, v, v_in and id are then read at
// permIn[idx[i]] when permIn is set (their builds were deferred).
struct Permute {
    const uint32_t* idx;
    const pbf_float4 *x, *p, *v, *vin;
    const uint32_t* id;
    pbf_float4 *xo, *po, *vo, *vino;
    uint32_t* ido;
    const uint32_t* permIn = nullptr;
    uint32_t* permOut = nullptr;
};
// launch_cell_ranges and the permutation in one pass.
pbf_status launch_ranges_gather(const uint32_t* sk, int* cs, int* ce, size_t n, const Permute& pm,
                                void* stream, uint32_t* splitters = nullptr);
size_t radix_scratch_words(size_t n);
// Scratch of a grid of capacity n: the radix sort's words, then the
// re-sort's counters (which must be zero before the first resort_pairs).
size_t grid_scratch_words(size_t n);
size_t resort_splitter_count(size_t n);
// Stable sort of (keys, slot) into kout/vout, equal to radix_sort_pairs',
// from prev, the previous sort's output keys of the same slots, and its
// splitters; fast when few keys changed.  kout may be prev; tmpK/tmpV are
// distinct from the rest; scratch holds grid_scratch_words(capacity).
pbf_status resort_pairs(const uint32_t* keys, const uint32_t* prev, const uint32_t* splitters,
                        uint32_t* kout, uint32_t* vout, uint32_t* tmpK, uint32_t* tmpV, size_t n,
                        size_t capacity, uint32_t* scratch, void* stream);
int radix_key_bits(size_t numCells);
// Stable sort of (key, value) pairs by the low `bits` key bits into
// kout/vout (vin == nullptr: values are the slots); the five arrays are
// distinct.
pbf_status radix_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                            uint32_t* vout, uint32_t* tmpK, uint32_t* tmpV, size_t n, int bits,
                            uint32_t* scratch, void* stream);

// Solver (k_solver.hip)
// With box not null, the keys are those of the container points of p*
// (skin lists); box is the single scene's container (a batched step uses
// the scene table's).  clearCells: -1 into the build's cellStart
// (clearCount cells).
pbf_status launch_predict_key(const pbf_float4* x, pbf_float4* p, pbf_float4* v, pbf_float4* vin,
                              uint32_t* key, int n, float dt, pbf_float3 g, pbf_float3 mn,
                              float cs, pbf_int3 gs, const pbf_container* box, float radius,
                              const SceneTable& st, int* clearCells, size_t clearCount,
                              const uint32_t* status, void* stream);
// Copies x, v and id to the frame checkpoint while *status is zero.
pbf_status launch_checkpoint(const pbf_float4* x, const pbf_float4* v, const uint32_t* id,
                             pbf_float4* cx, pbf_float4* cv, uint32_t* cid, size_t n,
                             const uint32_t* status, void* stream);
// Counts a finished frame in status[1] while status[0] is zero.
pbf_status launch_frame_done(uint32_t* status, void* stream);
pbf_status launch_keys(const pbf_float4* p, uint32_t* key, size_t n, pbf_float3 mn, float cs,
                       pbf_int3 gs, const SceneTable& st, void* stream);
// lambda goes to p[].w (the step's Delta-p reads it there).
pbf_status launch_lambda_guarded(pbf_float4* p, const uint32_t* nbr, const int* cnt, int maxN,
                                 size_t n, float h, float mass, float rho0, float eps,
                                 size_t stride, const ListPass& lp, uint32_t* status, void* stream);
// Reads lambda from pin[].w and restores pout[].w from x[].w.
pbf_status launch_delta_apply_collide(const pbf_float4* pin, pbf_float4* pout, const uint32_t* nbr,
                                      const int* cnt, int maxN, const pbf_float4* x, size_t n,
                                      float h, float mass, float rho0, float sk, int sn, float sdq,
                                      size_t stride, const pbfdev::Colliders& col, float radius,
                                      const SceneTable& st, const ListPass& lp, uint32_t* status,
                                      void* stream);
pbf_status launch_update_xsph(const pbf_float4* x, const pbf_float4* p, const pbf_float4* vold,
                              pbf_float4* vout, const uint32_t* nbr, const int* cnt, int maxN,
                              size_t n, float h, float visc, float inv_dt, size_t stride,
                              const ListPass& lp, uint32_t* status, void* stream);
pbf_status launch_update_velocity_only(const pbf_float4* x, const pbf_float4* p, pbf_float4* v,
                                       size_t n, float inv_dt, const uint32_t* status,
                                       void* stream);
// Writes omega (with |omega| in .w) and the records pw = (p.xyz, |omega|).
// lp.exactMax > 0: also counts the neighbours within h (the post-solve
// build's overflow check when XSPH is off).
pbf_status launch_vorticity(const pbf_float4* p, const pbf_float4* v, const uint32_t* nbr,
                            const int* cnt, int maxN, size_t n, float h, pbf_float4* omega,
                            pbf_float4* pw, size_t stride, const ListPass& lp, uint32_t* status,
                            void* stream);
pbf_status launch_confine_resolve(const pbf_float4* x, const pbf_float4* pw, const uint32_t* nbr,
                                  const int* cnt, int maxN, const pbf_float4* omega,
                                  const pbf_float4* vcur,
                                  const pbf_float4* vsnap, pbf_float4* vout, size_t n, float h,
                                  float dt, float strength, size_t stride,
                                  const pbfdev::Colliders& col, float radius, float e, float f,
                                  const SceneTable& st, const ListPass& lp,
                                  const uint32_t* status, void* stream);
pbf_status launch_resolve_guarded(const pbf_float4* x, const pbf_float4* vsnap, pbf_float4* v,
                                  size_t n, const pbfdev::Colliders& col, float r, float e,
                                  float f, const SceneTable& st, const uint32_t* status,
                                  void* stream);
pbf_status launch_zero_u32(uint32_t* p, size_t n, void* stream);

// Adjoint of the substep (k_adjoint.hip).  One grid build of the forward's
// tape: its grouped lists and counts, and the particle id of every slot.
struct AdjBuild {
    const uint32_t* nbr;
    const int* cnt;
    const uint32_t* ids;
    size_t stride;
};
// The post-solve build and the records of that phase (its slot order).
struct AdjPost {
    AdjBuild build;
    const pbf_float4* xs;     // substep start positions
    const pbf_float4* pf;     // final positions
    const pbf_float4* v3;     // velocities after XSPH
    const pbf_float4* omega;  // vorticity, |omega| in .w
    const pbf_float4* vin;    // velocities after gravity
};
// One constraint iteration: positions entering it (lambda in .w) and after
// its Delta-p, before the collision solve, in its build's slot order.
struct AdjIter {
    AdjBuild build;
    const pbf_float4* pk;
    const pbf_float4* pre;
};
// Adjoint state.  pb, vb (the adjoints of a substep's output positions and
// velocities), x0b, v1b (of its input positions and velocities) are in
// particle-id order; the rest are per-slot temporaries; acc holds the
// adjoints of the XSPH and vorticity strengths (FP64).
struct AdjBuffers {
    pbf_float4* pb;
    const pbf_float4* vb;
    pbf_float4* x0b;
    pbf_float4* v1b;
    pbf_float4* v3b;
    pbf_float4* etab;
    pbf_float4* wb;
    pbf_float4* prb;
    pbf_float4* gsum;
    float* cterm;
    float* lamb;
    double* acc;
};
struct AdjParams {
    float h, dt, visc, vort, radius, restitution, friction;
    float mass, rho0, lambdaEps, scorrK, scorrDq;
    int scorrN;
    pbfdev::Colliders col;  // the container, spheres, boxes and planes (device tables)
};
// Post-solve phase backward: from pb (output positions) and vb (output
// velocities) to x0b, v1b partially and pb (the last iteration's output).
pbf_status adjoint_post(const AdjPost& t, const AdjBuffers& w, size_t n, const AdjParams& p,
                        void* stream);
// One constraint iteration backward: pb from its output to its input.
pbf_status adjoint_iteration(const AdjIter& t, const AdjBuffers& w, size_t n,
                             const AdjParams& p, void* stream);
// Integrate backward: completes x0b and v1b from pb.
pbf_status adjoint_predict(const AdjBuffers& w, size_t n, float dt, void* stream);
// One grid build into the adjoint's tape: the groups of the grouped lists
// (stride slots per group row) that hold entries, ceil(cnt[i] / 4) of slot i
// (at most maxGroups), the counts and the slot-to-id map.
pbf_status tape_build(const uint32_t* lists, uint32_t* listsOut, const int* cnt, int* cntOut,
                      const uint32_t* ids, uint32_t* idsOut, size_t n, size_t stride,
                      int maxGroups, void* stream);
// Three record arrays of n records into the tape, in one launch.
pbf_status tape_records(const pbf_float4* a, pbf_float4* aOut, const pbf_float4* b,
                        pbf_float4* bOut, const pbf_float4* c, pbf_float4* cOut, size_t n,
                        void* stream);

// Neighbour lists by LDS-staged cell sweep (k_sweep.hip).  The grid must
// have cell size >= h (a 3x3x3 stencil) and the working set must be in the
// order of sortedKeys.
struct SweepGrid {
    const int* cellStart;
    const int* cellEnd;
    const uint32_t* sortedKeys;
    pbf_int3 dims;  // one scene's grid (unused when st is set)
    size_t n;
    float h;
    int maxNeighbors;
    SceneTable st;  // a batched step's scenes
    // Skin lists: positions are taken at their container points (box, or
    // the scene table's containers, shrunk by radius), and a list longer
    // than maxNeighbors raises PBF_STATUS_LIST_STALE instead of the
    // overflow bit.
    bool skin = false;
    const pbf_container* box = nullptr;
    float radius = 0.0f;
    // XSPH fused into the build (exact lists only): with vout set, every
    // particle's velocity after applyXsphViscosity (the arithmetic of
    // launch_update_xsph, its terms added in list order as the list is
    // built) into vout and, unless null, tapeOut; xs holds the start
    // positions (v = (p - xs) / dt), vold the velocities (.w kept).
    const pbf_float4* xs = nullptr;
    const pbf_float4* vold = nullptr;
    pbf_float4* vout = nullptr;
    pbf_float4* tapeOut = nullptr;
    float visc = 0.0f;
    float inv_dt = 0.0f;
};
// The lists launch_find_neighbors builds (at most maxNeighbors stored,
// counts whole), bit for bit, in grouped layout: entry q of particle i at
// word 4 * ((q / 4) * stride + i) + q % 4, and the slots of the last group
// past the count repeat its last entry (maxNeighbors a multiple of 4, nbr
// 16-byte aligned).  Returns at once when *status is non-zero and raises
// PBF_STATUS_NEIGHBOR_OVERFLOW.
pbf_status sweep_build_lists(const SweepGrid& g, const pbf_float4* p, uint32_t* nbr, int* cnt,
                             size_t stride, uint32_t* status, void* stream);
// CollisionSystem::solve as one launch: per-particle fixed point, at most 32
// passes; sets PBF_STATUS_COLLISION_NONCONVERGENCE in *status otherwise.
pbf_status launch_collide_fixed_point(pbf_float4* p, size_t n, const pbfdev::Colliders& col,
                                      float radius, uint32_t* status, void* stream);

}  // namespace pbfint
