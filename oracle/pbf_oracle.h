/*
 * pbf_oracle.h -- CPU restatement of the DusanCojic/Fluid-Simulator PBF path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as a checker or
 * as the timed CPU baseline.  The product library never links it.
 *
 * One source, two builds:
 *   liboracle_f64.so  ORC_REAL=double: every kernel's arithmetic in FP64,
 *                     results rounded to FP32 exactly where the reference
 *                     stores them (the parity oracle);
 *   liboracle_f32.so  ORC_REAL=float with OpenMP: the reference dataflow in
 *                     FP32 on the host cores (the CPU baseline).
 * In both builds the cell hash, the inclusive neighbour test and the
 * collision projections use the reference's FP32 arithmetic, because their
 * results are defined by FP32 rounding (grid indices, fixed points).
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

typedef struct orc_f3 { float x, y, z; } orc_f3;
typedef struct orc_f4 { float x, y, z, w; } __attribute__((aligned(16))) orc_f4;
typedef struct orc_i3 { int x, y, z; } orc_i3;

/* SimulationParams (include/pbf/simulation_params.hpp:6-29), 72 bytes. */
typedef struct orc_params {
    float dt, restDensity, particleMass, particleRadius, collisionRestitution,
        collisionFriction, smoothingRadius, lambdaEpsilon;
    int solverIterations, substeps;
    orc_f3 gravity;
    float scorrK;
    int scorrN;
    float scorrDeltaQ, xsphViscosity, vorticityStrength;
} orc_params;

typedef struct orc_container { orc_f3 min, max; } orc_container;
typedef struct orc_sphere { orc_f3 center; float radius; } orc_sphere;
typedef struct orc_box { orc_f3 center, half; } orc_box;
typedef struct orc_plane { orc_f3 point, normal; } orc_plane;

/* Status values equal the library's pbf_status codes. */
enum { ORC_OK = 0, ORC_INVALID = 1, ORC_LOGIC = 2, ORC_OUT_OF_RANGE = 3, ORC_LENGTH = 4,
       ORC_OVERFLOW = 5, ORC_NONCONVERGENCE = 6 };

const char* orc_precision(void);

/* SPH helpers (include/pbf/sph_helpers.cuh:13-95,
 * include/pbf/pbf_solver_kernels.cuh:24-47). */
float orc_poly6(orc_f3 d, float h);
void orc_spiky(orc_f3 d, float h, orc_f3* out);
float orc_scorr(float d2, float h, float k, int n, float dq);

/* Grid (src/pbf/spatial_grid.cu:10-137, include/pbf/spatial_grid.cuh:65-102). */
int orc_grid_dims(size_t maxParticles, orc_f3 mn, orc_f3 mx, float cellSize, orc_i3* dims,
                  size_t* cells);
void orc_cell_keys(const orc_f4* p, size_t n, orc_f3 mn, float cellSize, orc_i3 g,
                   uint32_t* keys);
void orc_grid_build(const orc_f4* p, size_t n, orc_f3 mn, float cellSize, orc_i3 g, size_t cells,
                    uint32_t* sortedKeys, uint32_t* sortedIdx, int* cellStart, int* cellEnd);

/* findNeighbors (src/pbf/neighbors.cu:3-91); returns 1 on overflow. */
int orc_find_neighbors(const orc_f4* p, const int* cellStart, const int* cellEnd, orc_i3 g,
                       orc_f3 mn, float cellSize, size_t n, float h, uint32_t* nbr, int* cnt,
                       int maxN, size_t stride);

/* Particle kernels (src/pbf/integration.cu, src/pbf/pbf_solver_kernels.cu). */
void orc_integrate(const orc_f4* x, orc_f4* p, orc_f4* v, size_t n, float dt, orc_f3 g);
void orc_lambda(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN, size_t n,
                float h, float mass, float rho0, float eps, float* lambda, size_t stride);
void orc_delta(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN,
               const float* lambda, size_t n, float h, float mass, float rho0, float k, int sn,
               float dq, orc_f4* delta, size_t stride);
void orc_apply_delta(orc_f4* p, const orc_f4* delta, size_t n);
void orc_update_velocity(orc_f4* x, const orc_f4* p, orc_f4* v, size_t n, float invDt);
void orc_xsph(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN, const orc_f4* vin,
              orc_f4* vout, size_t n, float h, float c, size_t stride);
void orc_vorticity(const orc_f4* p, const orc_f4* v, const uint32_t* nbr, const int* cnt,
                   int maxN, size_t n, float h, orc_f4* omega, size_t stride);
void orc_confine(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN,
                 const orc_f4* omega, const orc_f4* vin, orc_f4* vout, size_t n, float h,
                 float dt, float eps, size_t stride);

/* Collisions (src/pbf/collision_kernels.cu:7-371, collision_system.cu:133-218). */
void orc_solve_container(orc_f4* p, size_t n, const orc_container* c, float r, int* flag);
void orc_solve_spheres(orc_f4* p, size_t n, const orc_sphere* s, size_t ns, float r, int* flag);
void orc_solve_boxes(orc_f4* p, size_t n, const orc_box* b, size_t nb, float r, int* flag);
void orc_solve_planes(orc_f4* p, size_t n, const orc_plane* pl, size_t np, float r, int* flag);
/* CollisionSystem::solve: whole passes until none marks, at most 32. */
int orc_collision_solve(orc_f4* p, size_t n, const orc_container* c, const orc_sphere* s,
                        size_t ns, const orc_box* b, size_t nb, const orc_plane* pl, size_t np,
                        float r);
void orc_resolve_velocities(const orc_f4* x, const orc_f4* vin, orc_f4* v, size_t n,
                            const orc_container* c, const orc_sphere* s, size_t ns,
                            const orc_box* b, size_t nb, const orc_plane* pl, size_t np, float r,
                            float e, float f);

/* PBFSolver (src/pbf/pbf_solver.cu:23-379). */
typedef struct orc_solver orc_solver;
int orc_solver_create(orc_solver** out, size_t maxParticles, orc_f3 mn, orc_f3 mx,
                      const orc_params* params);
void orc_solver_destroy(orc_solver* s);
int orc_solver_set_particles(orc_solver* s, const orc_f4* pos, const orc_f4* vel, size_t n);
int orc_solver_set_colliders(orc_solver* s, const orc_sphere* sp, size_t ns, const orc_box* b,
                             size_t nb, const orc_plane* pl, size_t np);
int orc_solver_step(orc_solver* s);
int orc_solver_run(orc_solver* s, size_t frames);
/* Positions and velocities in API (stable id) order. */
int orc_solver_get(const orc_solver* s, orc_f4* pos, orc_f4* vel, size_t n);
/* Working-set view (cell order) for bit-level checks: p*, stable ids. */
int orc_solver_working_set(const orc_solver* s, orc_f4* predicted, uint32_t* ids, size_t n);

#endif /* PBF_ORACLE_H */
