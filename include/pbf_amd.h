/*
 * pbf_amd.h -- C ABI of the MI355X (gfx950) Position Based Fluids library.
 *
 * The library replaces the CUDA library `pbf` of DusanCojic/Fluid-Simulator.
 * Every entry point below names the reference interface it stands in for.
 * The ABI is plain C: POD structs, raw device pointers, sizes, an opaque
 * stream handle and integer status codes.  No HIP, CUDA or torch type
 * appears in a signature.
 *
 * Conventions
 *  - Functions return a pbf_status.  On failure, pbf_last_error() returns a
 *    thread-local message.  The reference throws C++ exceptions; the status
 *    codes map one-to-one onto its exception types (see pbf_status).
 *  - Pointers named d_* or documented as "device" are device (HBM) pointers.
 *    Pointers named h_* are host pointers.
 *  - `stream` is a hipStream_t passed as void*; NULL means the legacy default
 *    stream, which is what the reference launches on.
 *  - Kernel launchers (pbf_k_*) are asynchronous: they return after the
 *    launch, like a CUDA `<<<>>>` launch followed by cudaGetLastError().
 *  - Object handles (pbf_grid, pbf_collision, pbf_solver) own device memory
 *    and synchronise where the reference does.
 */
#ifndef PBF_AMD_H
#define PBF_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Plain data types.  Byte layouts equal CUDA's float3/float4/int3 and the    */
/* reference PODs, so buffers can be shared with reference-shaped callers.   */
/* ------------------------------------------------------------------------ */

/* CUDA float3: 12 bytes, 4-byte aligned. */
typedef struct pbf_float3 { float x, y, z; } pbf_float3;

/* CUDA float4: 16 bytes, 16-byte aligned.  .w is carried, never interpreted. */
typedef struct pbf_float4 { float x, y, z, w; } __attribute__((aligned(16))) pbf_float4;

/* CUDA int3: 12 bytes, 4-byte aligned. */
typedef struct pbf_int3 { int x, y, z; } pbf_int3;

/* Replaces SimulationParams (include/pbf/simulation_params.hpp:6-29).
 * 72 bytes; the field order and offsets are those of the reference. */
typedef struct pbf_sim_params {
    float dt;                   /*  0: frame step; a substep is dt / substeps */
    float restDensity;          /*  4 */
    float particleMass;         /*  8 */
    float particleRadius;       /* 12: collision radius */
    float collisionRestitution; /* 16 */
    float collisionFriction;    /* 20 */
    float smoothingRadius;      /* 24: SPH kernel radius h */
    float lambdaEpsilon;        /* 28: constraint relaxation */
    int   solverIterations;     /* 32 */
    int   substeps;             /* 36 */
    pbf_float3 gravity;         /* 40 */
    float scorrK;               /* 52: artificial pressure strength */
    int   scorrN;               /* 56: artificial pressure exponent */
    float scorrDeltaQ;          /* 60: artificial pressure reference distance */
    float xsphViscosity;        /* 64 */
    float vorticityStrength;    /* 68 */
} pbf_sim_params;

/* Replace Container / SphereCollider / BoxCollider / PlaneCollider
 * (include/pbf/colliders.hpp:7-25). */
typedef struct pbf_container { pbf_float3 min, max; } pbf_container;          /* 24 B */
typedef struct pbf_sphere { pbf_float3 center; float radius; } pbf_sphere;    /* 16 B */
typedef struct pbf_box { pbf_float3 center, halfExtents; } pbf_box;          /* 24 B */
typedef struct pbf_plane { pbf_float3 point, normal; } pbf_plane;            /* 24 B */

/* Status codes.  The right-hand column is the exception the reference
 * throws for the same condition. */
typedef enum pbf_status {
    PBF_OK = 0,
    PBF_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    PBF_ERR_LOGIC = 2,            /* std::logic_error (use before initialize) */
    PBF_ERR_OUT_OF_RANGE = 3,     /* std::out_of_range */
    PBF_ERR_LENGTH = 4,           /* std::length_error */
    PBF_ERR_OVERFLOW = 5,         /* std::overflow_error (neighbour capacity) */
    PBF_ERR_NONCONVERGENCE = 6,   /* std::runtime_error (collision passes) */
    PBF_ERR_DEVICE = 7            /* std::runtime_error (HIP runtime error) */
} pbf_status;

/* Bits of the device status word the fused step writes instead of the
 * reference's per-phase host synchronisations. */
#define PBF_STATUS_NEIGHBOR_OVERFLOW 0x1u
#define PBF_STATUS_COLLISION_NONCONVERGENCE 0x2u
/* A particle moved farther than half the skin of the step's skin lists
 * (pbf_solver_set_neighbor_skin), or a skin list ran out of room: the
 * frame is run again with a grid build per iteration.  Never an error. */
#define PBF_STATUS_LIST_STALE 0x4u

/* Thread-local message for the last failing call on this thread. */
const char* pbf_last_error(void);

/* Library version string ("pbf_amd <version> gfx950"). */
const char* pbf_version(void);

/* ------------------------------------------------------------------------ */
/* Device memory helpers (stand in for CudaBuffer<T>,                         */
/* include/pbf/cuda_buffer.hpp:43-113).                                       */
/* ------------------------------------------------------------------------ */
pbf_status pbf_device_alloc(void** d_ptr, size_t bytes);
pbf_status pbf_device_free(void* d_ptr);
pbf_status pbf_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
pbf_status pbf_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
pbf_status pbf_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
/* Memset semantics: every byte gets the low 8 bits of byteValue
 * (CudaBuffer::fillBytes, cuda_buffer.hpp:67-77). */
pbf_status pbf_memset(void* d_dst, int byteValue, size_t bytes);
/* Pinned (page-locked) host memory.  No counterpart in the reference: the
 * solver's uploads from and copy-outs into such memory run as one DMA each
 * instead of going through the solver's pinned staging chunks. */
pbf_status pbf_host_alloc(void** h_ptr, size_t bytes);
pbf_status pbf_host_free(void* h_ptr);
pbf_status pbf_device_synchronize(void);
pbf_status pbf_stream_create(void** stream);
pbf_status pbf_stream_destroy(void* stream);
pbf_status pbf_stream_synchronize(void* stream);

/* ------------------------------------------------------------------------ */
/* Kernel launchers.  Each replaces one __global__ kernel of the reference   */
/* and keeps its argument meaning.  Launches are asynchronous on `stream`.   */
/* ------------------------------------------------------------------------ */

/* integrate (include/pbf/integration.cuh:6-8, src/pbf/integration.cu:6-32):
 * v += g*dt; p* = x + v*dt; p*.w = x.w. */
pbf_status pbf_k_integrate(const pbf_float4* d_position, pbf_float4* d_predicted,
                           pbf_float4* d_velocity, int particleCount, float dt,
                           pbf_float3 gravity, void* stream);

/* findNeighbors (include/pbf/neighbors.cuh:16-23, src/pbf/neighbors.cu:3-91).
 * Slot-major list: d_neighbors[k * particleStride + i].  d_count receives the
 * full count, which may exceed maxNeighbors; then *d_overflowFlag = 1 if the
 * flag pointer is not NULL. */
pbf_status pbf_k_find_neighbors(const pbf_float4* d_predicted, const int* d_cellStart,
                                const int* d_cellEnd, pbf_int3 gridSize,
                                pbf_float3 minBounds, float cellSize,
                                size_t particleCount, float smoothingRadius,
                                uint32_t* d_neighbors, int* d_count, int maxNeighbors,
                                size_t particleStride, int* d_overflowFlag, void* stream);

/* computeLambda (include/pbf/pbf_solver_kernels.cuh:16-21,
 * src/pbf/pbf_solver_kernels.cu:26-105). */
pbf_status pbf_k_compute_lambda(const pbf_float4* d_predicted, const uint32_t* d_neighbors,
                                const int* d_count, int maxNeighbors, size_t particleCount,
                                float smoothingRadius, float particleMass,
                                float restDensity, float epsilon, float* d_lambda,
                                size_t neighborParticleStride, void* stream);

/* computeDeltaPosition (include/pbf/pbf_solver_kernels.cuh:49-53,
 * src/pbf/pbf_solver_kernels.cu:108-182), with computeArtificialPressure
 * (include/pbf/pbf_solver_kernels.cuh:24-47). */
pbf_status pbf_k_compute_delta_position(const pbf_float4* d_predicted,
                                        const uint32_t* d_neighbors, const int* d_count,
                                        int maxNeighbors, const float* d_lambda,
                                        size_t particleCount, float smoothingRadius,
                                        float particleMass, float restDensity, float scorrK,
                                        int scorrN, float scorrDeltaQ, pbf_float4* d_delta,
                                        size_t neighborParticleStride, void* stream);

/* applyDeltaPosition (src/pbf/pbf_solver_kernels.cu:184-193). */
pbf_status pbf_k_apply_delta_position(pbf_float4* d_predicted, const pbf_float4* d_delta,
                                      size_t particleCount, void* stream);

/* updateVelocityAndPosition (src/pbf/pbf_solver_kernels.cu:195-211). */
pbf_status pbf_k_update_velocity_and_position(pbf_float4* d_position,
                                              const pbf_float4* d_predicted,
                                              pbf_float4* d_velocity, size_t particleCount,
                                              float inverseDt, void* stream);

/* applyXsphViscosity (include/pbf/pbf_solver_kernels.cuh:63-70,
 * src/pbf/pbf_solver_kernels.cu:213-266).  Jacobi: reads d_vin, writes d_vout. */
pbf_status pbf_k_apply_xsph_viscosity(const pbf_float4* d_predicted,
                                      const uint32_t* d_neighbors, const int* d_count,
                                      int maxNeighbors, const pbf_float4* d_vin,
                                      pbf_float4* d_vout, size_t particleCount,
                                      float smoothingRadius, float xsphViscosity,
                                      size_t neighborParticleStride, void* stream);

/* computeVorticity (include/pbf/pbf_solver_kernels.cuh:72-79,
 * src/pbf/pbf_solver_kernels.cu:268-315). */
pbf_status pbf_k_compute_vorticity(const pbf_float4* d_position, const pbf_float4* d_velocity,
                                   const uint32_t* d_neighbors, const int* d_count,
                                   int maxNeighbors, size_t particleCount,
                                   float smoothingRadius, pbf_float4* d_vorticity,
                                   size_t neighborParticleStride, void* stream);

/* applyVorticityConfinement (include/pbf/pbf_solver_kernels.cuh:81-93,
 * src/pbf/pbf_solver_kernels.cu:317-385). */
pbf_status pbf_k_apply_vorticity_confinement(const pbf_float4* d_position,
                                             const uint32_t* d_neighbors, const int* d_count,
                                             int maxNeighbors, const pbf_float4* d_vorticity,
                                             const pbf_float4* d_vin, pbf_float4* d_vout,
                                             size_t particleCount, float smoothingRadius,
                                             float dt, float vorticityStrength,
                                             size_t neighborParticleStride, void* stream);

/* gatherSolverState (include/pbf/pbf_solver_kernels.cuh:95-105,
 * src/pbf/pbf_solver_kernels.cu:387-407). */
pbf_status pbf_k_gather_solver_state(const uint32_t* d_sortedIndices,
                                     const pbf_float4* d_position, const pbf_float4* d_predicted,
                                     const pbf_float4* d_velocity,
                                     const pbf_float4* d_collisionInputVelocity,
                                     const uint32_t* d_stableIds, pbf_float4* d_sortedPosition,
                                     pbf_float4* d_sortedPredicted, pbf_float4* d_sortedVelocity,
                                     pbf_float4* d_sortedCollisionInputVelocity,
                                     uint32_t* d_sortedStableIds, size_t particleCount,
                                     void* stream);

/* initializeStableParticleIds (src/pbf/pbf_solver_kernels.cu:409-416). */
pbf_status pbf_k_initialize_stable_particle_ids(uint32_t* d_ids, size_t particleCount,
                                                void* stream);

/* scatterFloat4ByStableId (src/pbf/pbf_solver_kernels.cu:418-427). */
pbf_status pbf_k_scatter_float4_by_stable_id(const pbf_float4* d_values, const uint32_t* d_ids,
                                             pbf_float4* d_valuesInOriginalOrder,
                                             size_t particleCount, void* stream);

/* Collision kernels (include/pbf/collision_kernels.cuh:8-34,
 * src/pbf/collision_kernels.cu:47-371).  Collider pointers are device
 * pointers; d_correctionFlag may be NULL. */
pbf_status pbf_k_solve_container(pbf_float4* d_predicted, size_t particleCount,
                                 const pbf_container* d_container, float particleRadius,
                                 int* d_correctionFlag, void* stream);
pbf_status pbf_k_solve_spheres(pbf_float4* d_predicted, size_t particleCount,
                               const pbf_sphere* d_spheres, size_t sphereCount,
                               float particleRadius, int* d_correctionFlag, void* stream);
pbf_status pbf_k_solve_boxes(pbf_float4* d_predicted, size_t particleCount,
                             const pbf_box* d_boxes, size_t boxCount, float particleRadius,
                             int* d_correctionFlag, void* stream);
pbf_status pbf_k_solve_planes(pbf_float4* d_predicted, size_t particleCount,
                              const pbf_plane* d_planes, size_t planeCount,
                              float particleRadius, int* d_correctionFlag, void* stream);
pbf_status pbf_k_resolve_velocities(const pbf_float4* d_position,
                                    const pbf_float4* d_incomingVelocity,
                                    pbf_float4* d_velocity, size_t particleCount,
                                    const pbf_container* d_container,
                                    const pbf_sphere* d_spheres, size_t sphereCount,
                                    const pbf_box* d_boxes, size_t boxCount,
                                    const pbf_plane* d_planes, size_t planeCount,
                                    float particleRadius, float restitution, float friction,
                                    void* stream);

/* Device evaluation of the inline SPH helpers, batched (the reference tests
 * reach them through single-thread test kernels, test/test_sph_helpers.cu:10-18
 * and test/test_pbf_solver_kernels.cu:14-26):
 * poly6 (include/pbf/sph_helpers.cuh:38-46), spikyGradient (:85-95),
 * computeArtificialPressure (include/pbf/pbf_solver_kernels.cuh:24-47). */
pbf_status pbf_k_poly6(const pbf_float3* d_displacement, size_t count, float smoothingRadius,
                       float* d_out, void* stream);
pbf_status pbf_k_spiky_gradient(const pbf_float3* d_displacement, size_t count,
                                float smoothingRadius, pbf_float3* d_out, void* stream);
pbf_status pbf_k_artificial_pressure(const float* d_distanceSquared, size_t count,
                                     float smoothingRadius, float scorrK, int scorrN,
                                     float scorrDeltaQ, float* d_out, void* stream);

/* ------------------------------------------------------------------------ */
/* SpatialGrid (include/pbf/spatial_grid.cuh:11-62, src/pbf/spatial_grid.cu). */
/* Cell size, bounds and the 32-bit key layout x + y*gx + z*gx*gy are the    */
/* reference's; the sort is this library's LDS-staged LSD radix sort over    */
/* ceil(log2(numCells)) key bits, stable in the input slot.                  */
/* ------------------------------------------------------------------------ */
typedef struct pbf_grid pbf_grid;

pbf_status pbf_grid_create(pbf_grid** grid);
pbf_status pbf_grid_destroy(pbf_grid* grid);
/* SpatialGrid::initialize (spatial_grid.cu:10-97).  Strong guarantee. */
pbf_status pbf_grid_initialize(pbf_grid* grid, size_t maxParticles, pbf_float3 minBounds,
                               pbf_float3 maxBounds, float cellSize);
/* SpatialGrid::build (spatial_grid.cu:139-226): keys, stable sort, ranges.
 * Synchronises before returning, like the reference. */
pbf_status pbf_grid_build(pbf_grid* grid, const pbf_float4* d_predicted, size_t particleCount);
/* pbf_grid_build's results bit for bit, computed by re-sorting from the
 * grid's last build.  No counterpart in the reference: it is for callers
 * that lay their particles out in the last build's order (slot i holding
 * particle pbf_grid_sorted_indices()[i], as the solver's reorder does) and
 * build again after a small move; then only the keys that changed are
 * sorted and merged into the kept ones.  Correct for any positions; without
 * a previous build of particleCount particles it is pbf_grid_build.
 * Synchronises before returning. */
pbf_status pbf_grid_rebuild(pbf_grid* grid, const pbf_float4* d_predicted, size_t particleCount);
/* Entry point that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is
 * pbf_grid_build without a grid handle: the grid is given by minBounds,
 * maxBounds and cellSize (validated like pbf_grid_initialize), the outputs
 * are the caller's (d_sortedKeys and d_sortedIndices hold particleCount
 * entries, d_cellStart and d_cellEnd cellCapacity >= the grid's cells), and
 * the keys, the sort's ping-pong buffers and its counters live in one
 * static device array of about 13.6 MB per GPU, held until the process
 * ends.  Results equal pbf_grid_build's bit for bit.  At most
 * PBF_GRID_BUILD_STATIC_MAX_PARTICLES particles (PBF_ERR_LENGTH above).
 * Asynchronous on `stream` (no host synchronisation, no memset node).
 * Calls on one GPU share its array: calls that overlap on one GPU (other
 * streams or threads) race; calls ordered on one stream do not. */
#define PBF_GRID_BUILD_STATIC_MAX_PARTICLES 1048576u
pbf_status pbf_grid_build_static(const pbf_float4* d_predicted, size_t particleCount,
                                 pbf_float3 minBounds, pbf_float3 maxBounds, float cellSize,
                                 uint32_t* d_sortedKeys, uint32_t* d_sortedIndices,
                                 int* d_cellStart, int* d_cellEnd, size_t cellCapacity,
                                 void* stream);
/* Device-pointer getters (spatial_grid.cu:230-244); valid until the next
 * initialize/build. */
const uint32_t* pbf_grid_sorted_keys(const pbf_grid* grid);
const uint32_t* pbf_grid_sorted_indices(const pbf_grid* grid);
const int* pbf_grid_cell_start(const pbf_grid* grid);
const int* pbf_grid_cell_end(const pbf_grid* grid);
/* Value getters (spatial_grid.cu:246-268).  Any out pointer may be NULL. */
pbf_status pbf_grid_info(const pbf_grid* grid, pbf_int3* gridSize, size_t* numCells,
                         size_t* maxParticles, float* cellSize, pbf_float3* minBounds,
                         pbf_float3* maxBounds);

/* ------------------------------------------------------------------------ */
/* CollisionSystem (include/pbf/collision_system.hpp:11-57,                  */
/* src/pbf/collision_system.cu:29-277).                                      */
/* ------------------------------------------------------------------------ */
typedef struct pbf_collision pbf_collision;

pbf_status pbf_collision_create(pbf_collision** cs);
pbf_status pbf_collision_destroy(pbf_collision* cs);
pbf_status pbf_collision_set_container(pbf_collision* cs, pbf_container container);
/* Host arrays; count 0 clears (collision_system.cu:52-131). */
pbf_status pbf_collision_set_spheres(pbf_collision* cs, const pbf_sphere* h_spheres, size_t count);
pbf_status pbf_collision_clear_spheres(pbf_collision* cs);
pbf_status pbf_collision_set_boxes(pbf_collision* cs, const pbf_box* h_boxes, size_t count);
pbf_status pbf_collision_clear_boxes(pbf_collision* cs);
pbf_status pbf_collision_set_planes(pbf_collision* cs, const pbf_plane* h_planes, size_t count);
pbf_status pbf_collision_clear_planes(pbf_collision* cs);
/* CollisionSystem::solve (collision_system.cu:133-172): projection to a fixed
 * point, at most 32 passes; PBF_ERR_NONCONVERGENCE otherwise.  Synchronous. */
pbf_status pbf_collision_solve(pbf_collision* cs, pbf_float4* d_predicted, size_t particleCount,
                               float particleRadius);
/* CollisionSystem::resolveVelocities, both overloads (collision_system.cu:174-218);
 * pass d_incomingVelocity == d_velocity for the two-pointer form. */
pbf_status pbf_collision_resolve_velocities(pbf_collision* cs, const pbf_float4* d_position,
                                            const pbf_float4* d_incomingVelocity,
                                            pbf_float4* d_velocity, size_t particleCount,
                                            float particleRadius, float restitution,
                                            float friction);

/* ------------------------------------------------------------------------ */
/* PBFSolver (include/pbf/pbf_solver.hpp:13-70, src/pbf/pbf_solver.cu).       */
/* ------------------------------------------------------------------------ */
typedef struct pbf_solver pbf_solver;

pbf_status pbf_solver_create(pbf_solver** solver);
pbf_status pbf_solver_destroy(pbf_solver* solver);
/* PBFSolver::initialize (pbf_solver.cu:23-116). */
pbf_status pbf_solver_initialize(pbf_solver* solver, size_t maxParticles, pbf_float3 minBounds,
                                 pbf_float3 maxBounds, const pbf_sim_params* params);
/* PBFSolver::setParticles (pbf_solver.cu:118-152): host arrays, validated
 * before any upload. */
pbf_status pbf_solver_set_particles(pbf_solver* solver, const pbf_float4* h_positions,
                                    const pbf_float4* h_velocities, size_t particleCount);
/* Same contract with device arrays (no host validation of values; the
 * caller guarantees finiteness).  Used when inputs are already in HBM. */
pbf_status pbf_solver_set_particles_device(pbf_solver* solver, const pbf_float4* d_positions,
                                           const pbf_float4* d_velocities, size_t particleCount);
/* PBFSolver::setSpheres/setBoxes/setPlanes (pbf_solver.cu:174-184). */
pbf_status pbf_solver_set_spheres(pbf_solver* solver, const pbf_sphere* h_spheres, size_t count);
pbf_status pbf_solver_set_boxes(pbf_solver* solver, const pbf_box* h_boxes, size_t count);
pbf_status pbf_solver_set_planes(pbf_solver* solver, const pbf_plane* h_planes, size_t count);
/* PBFSolver::step (pbf_solver.cu:186-328): one frame of params.substeps
 * substeps.  Returns after the device finished, like the reference.
 * Overflow and non-convergence are detected on the device and reported here
 * as PBF_ERR_OVERFLOW / PBF_ERR_NONCONVERGENCE. */
pbf_status pbf_solver_step(pbf_solver* solver);
/* PBFSolver::run() (pbf_solver.cu:330-335). */
pbf_status pbf_solver_run_one(pbf_solver* solver);
/* PBFSolver::run(frameCount) (pbf_solver.cu:337-343).  Frames are enqueued
 * back to back and the status word is read once at the end. */
pbf_status pbf_solver_run(pbf_solver* solver, size_t frameCount);
/* PBFSolver::copyPositionsToHost / copyVelocitiesToHost (pbf_solver.cu:345-379):
 * API (stable id) order. */
pbf_status pbf_solver_copy_positions_to_host(pbf_solver* solver, pbf_float4* h_out,
                                             size_t particleCount);
pbf_status pbf_solver_copy_velocities_to_host(pbf_solver* solver, pbf_float4* h_out,
                                              size_t particleCount);
/* Device-output variants of the two copies (API order into d_out). */
pbf_status pbf_solver_copy_positions_to_device(pbf_solver* solver, pbf_float4* d_out,
                                               size_t particleCount);
pbf_status pbf_solver_copy_velocities_to_device(pbf_solver* solver, pbf_float4* d_out,
                                                size_t particleCount);
/* Active particle count and capacity. */
pbf_status pbf_solver_counts(const pbf_solver* solver, size_t* particleCount,
                             size_t* maxParticles);
/* The stream the solver enqueues on (hipStream_t as void*). */
void* pbf_solver_stream(pbf_solver* solver);

/* Frames of pbf_solver_step / pbf_solver_run replayed from HIP graphs
 * (enabled = 1) or launched kernel by kernel (0, the default).  No
 * counterpart in the reference, whose step() launches every kernel from the
 * host and synchronises about 15 times: the status word already removed the
 * syncs, and a graph per frame removes the per-launch host work.  Results
 * are bitwise the same either way.  Graphs are captured on first use and
 * dropped by initialize, set_particles and the collider setters; frames run
 * eagerly while profiling is on. */
pbf_status pbf_solver_set_graphs(pbf_solver* solver, int enabled);

/* Grid builds by re-sort (enabled = 1, the default) or by full radix sort
 * (0).  No counterpart in the reference, which sorts every build's keys in
 * full (spatial_grid.cu:205-213).  The step's builds find the particles in
 * the order of the previous build's sort, and only a few per cent of their
 * cell keys change from one build to the next: those are sorted on their
 * own and merged into the kept ones (pbf_grid_rebuild).  Results are
 * bitwise the same either way; the option exists to compare the two. */
pbf_status pbf_solver_set_incremental_sort(pbf_solver* solver, int enabled);

/* XSPH fused into the post-solve grid build (enabled = 1, the default) or
 * run as a pass of its own after it (0).  No counterpart in the reference,
 * which runs applyXsphViscosity over the rebuilt lists
 * (pbf_solver.cu:284-298).  The fused build adds each neighbour's XSPH term
 * as it appends the neighbour to the list, with the arithmetic of the
 * separate pass and in list order, so results are bitwise the same either
 * way; the option exists to compare the two. */
pbf_status pbf_solver_set_fused_xsph(pbf_solver* solver, int enabled);

/* Skin lists (Verlet lists), no counterpart in the reference.  With skin > 0
 * a substep builds the grid and the neighbour lists once, instead of before
 * every constraint iteration and after the solve (pbf_solver.cu:210-282):
 * around the container points of the predicted positions, with radius
 * h * (1 + skin) on cells of that size.  Every pass still tests each entry
 * against h with the reference's FP32 test, so it sums the same neighbours
 * (in the skin list's order), and the counts within h raise the reference's
 * overflow_error.  The lists are complete while no particle has moved more
 * than 0.45 * skin * h from its container point: a Delta-p pass that sees a
 * particle past that marks the lists stale, and the frame runs again from a
 * checkpoint taken at its start with the reference's builds (counted by
 * staleFrames).  skin lies in [0, 1]; 0 (the default) runs the reference's
 * builds.  Results equal the reference's within the FP32 rounding of the
 * summation order. */
pbf_status pbf_solver_set_neighbor_skin(pbf_solver* solver, float skin);
/* The skin, and the frames run again since initialize because their skin
 * lists went stale (either pointer may be null). */
pbf_status pbf_solver_neighbor_skin(const pbf_solver* solver, float* skin, size_t* staleFrames);

/* Launch tunings and the autotuner, no counterpart in the reference.  A
 * tuning is the skin (the grid's cell-size multiplier is 1 + neighborSkin,
 * see pbf_solver_set_neighbor_skin), the workgroup size of the neighbour
 * passes (64, 128, 192 or 256 threads) and the particles each of their
 * threads takes in turn (1, 2 or 4); results depend on neither of the
 * last two. */
typedef struct pbf_tuning {
    float neighborSkin;
    int passBlock;
    int particlesPerThread;
} pbf_tuning;
/* Size of the default candidate set: skins 0, 0.1, 0.2, 0.3 by workgroups
 * of 64, 128 and 256 threads by 1 or 2 particles per thread. */
#define PBF_DEFAULT_TUNING_COUNT 24
pbf_status pbf_solver_set_tuning(pbf_solver* solver, const pbf_tuning* tuning);
pbf_status pbf_solver_get_tuning(const pbf_solver* solver, pbf_tuning* tuning);
/* Runs `frames` frames of each candidate (candidates == NULL: the default
 * set) from the current state, restoring it after each; times them on the
 * host clock, stale frames run again included; applies the fastest, writes
 * it to *chosen and each candidate's milliseconds per frame to msPerFrame
 * (count entries, or PBF_DEFAULT_TUNING_COUNT for the default set; either
 * pointer may be null).  The particle state is left as it was. */
pbf_status pbf_solver_autotune(pbf_solver* solver, size_t frames, const pbf_tuning* candidates,
                               size_t count, double* msPerFrame, pbf_tuning* chosen);

/* Per-kernel timing on the solver's stream with HIP events.  While enabled,
 * every launch of the step is bracketed by events; pbf_solver_kernel_times
 * returns accumulated milliseconds and launch counts per kernel name.
 * Names are written as one '\n'-separated string. */
pbf_status pbf_solver_set_profiling(pbf_solver* solver, int enabled);
pbf_status pbf_solver_kernel_times(pbf_solver* solver, char* names, size_t namesBytes,
                                   double* milliseconds, uint64_t* launches, size_t capacity,
                                   size_t* kernelCount);

/* Introspection, no counterpart in the reference: the neighbour lists of the
 * solver's last grid build, in the reference's slot-major layout
 * (h_neighbors[q * particleCount + i], q < min(h_counts[i], maxNeighbors)),
 * together with the working-set positions they were built on (working-set
 * order, not API order).  After a frame with XSPH or vorticity enabled the
 * last build is the post-solve rebuild, whose positions are the current
 * ones.  particleCount must equal the active count; maxNeighbors caps the
 * slots copied (h_neighbors holds maxNeighbors * particleCount entries). */
pbf_status pbf_solver_neighbor_lists(pbf_solver* solver, pbf_float4* h_positions, int* h_counts,
                                     uint32_t* h_neighbors, size_t particleCount,
                                     int maxNeighbors);

/* ------------------------------------------------------------------------ */
/* Batched scenes (PBFBatchSolver, include/pbf/pbf_batch.hpp).  No          */
/* counterpart in the reference, whose callers loop PBFSolver over scenes:  */
/* many independent scenes, each with its own grid and container (its      */
/* bounds) and shared parameters, advanced by one launch per stage.  Each   */
/* scene's result equals, bit for bit, a pbf_solver holding it alone.       */
/* ------------------------------------------------------------------------ */
typedef struct pbf_batch pbf_batch;
typedef struct pbf_batch_scene {
    size_t particleCount;
    pbf_float3 minBounds;
    pbf_float3 maxBounds;
} pbf_batch_scene;

pbf_status pbf_batch_create(pbf_batch** batch);
pbf_status pbf_batch_destroy(pbf_batch* batch);
/* Validates every scene as pbf_solver_initialize would and allocates. */
pbf_status pbf_batch_initialize(pbf_batch* batch, const pbf_batch_scene* scenes,
                                size_t sceneCount, const pbf_sim_params* params);
/* All particles, scene after scene; particleCount = the scenes' total. */
pbf_status pbf_batch_set_particles(pbf_batch* batch, const pbf_float4* h_positions,
                                   const pbf_float4* h_velocities, size_t particleCount);
pbf_status pbf_batch_step(pbf_batch* batch);
pbf_status pbf_batch_run(pbf_batch* batch, size_t frameCount);
/* All particles, in the order of pbf_batch_set_particles. */
pbf_status pbf_batch_copy_positions_to_host(pbf_batch* batch, pbf_float4* h_out,
                                            size_t particleCount);
pbf_status pbf_batch_copy_velocities_to_host(pbf_batch* batch, pbf_float4* h_out,
                                             size_t particleCount);
pbf_status pbf_batch_counts(const pbf_batch* batch, size_t* sceneCount, size_t* particleCount);
pbf_status pbf_batch_set_graphs(pbf_batch* batch, int enabled);
pbf_status pbf_batch_set_incremental_sort(pbf_batch* batch, int enabled);
pbf_status pbf_batch_set_fused_xsph(pbf_batch* batch, int enabled);
/* Skin lists for every scene of the batch (pbf_solver_set_neighbor_skin).
 * A frame in which any scene's lists go stale runs again for all scenes
 * with the reference's builds. */
pbf_status pbf_batch_set_neighbor_skin(pbf_batch* batch, float skin);
pbf_status pbf_batch_stale_frames(const pbf_batch* batch, size_t* staleFrames);
/* Tunings and the autotuner over the whole batch (pbf_solver_autotune). */
pbf_status pbf_batch_set_tuning(pbf_batch* batch, const pbf_tuning* tuning);
pbf_status pbf_batch_autotune(pbf_batch* batch, size_t frames, const pbf_tuning* candidates,
                              size_t count, double* msPerFrame, pbf_tuning* chosen);
pbf_status pbf_batch_set_profiling(pbf_batch* batch, int enabled);
pbf_status pbf_batch_kernel_times(pbf_batch* batch, char* names, size_t namesBytes,
                                  double* milliseconds, uint64_t* launches, size_t capacity,
                                  size_t* kernelCount);

/* ------------------------------------------------------------------------ */
/* Adjoint (PBFAdjoint, include/pbf/pbf_adjoint.hpp).  No counterpart in the */
/* reference: reverse-mode derivatives of a solver's frames, with neighbour */
/* lists and the forward's discrete decisions held fixed, checkpointed      */
/* every checkpointEvery frames and replayed with a tape per segment.       */
/* Container, spheres, rounded boxes and planes; XSPH viscosity and        */
/* vorticity strength may be zero (see include/pbf/pbf_adjoint.hpp).       */
/* ------------------------------------------------------------------------ */
typedef struct pbf_adjoint pbf_adjoint;

/* The solver must outlive the adjoint. */
pbf_status pbf_adjoint_create(pbf_solver* solver, pbf_adjoint** adjoint);
pbf_status pbf_adjoint_destroy(pbf_adjoint* adjoint);
/* Runs frameCount frames of the solver, keeping checkpoints. */
pbf_status pbf_adjoint_forward(pbf_adjoint* adjoint, size_t frameCount, size_t checkpointEvery);
/* dL/dx, dL/dv of the final state in (host, API order; NULL = zero), dL/dx,
 * dL/dv of the initial state and dL/d(xsphViscosity),
 * dL/d(vorticityStrength) out (any may be NULL).  The solver is left in the
 * state pbf_adjoint_forward ended in. */
pbf_status pbf_adjoint_backward(pbf_adjoint* adjoint, const pbf_float4* h_xBar,
                                const pbf_float4* h_vBar, size_t particleCount,
                                pbf_float4* h_x0Bar, pbf_float4* h_v0Bar, double* viscosityBar,
                                double* vorticityBar);

#ifdef __cplusplus
}
#endif

#endif /* PBF_AMD_H */
