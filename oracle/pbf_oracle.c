/*
 * pbf_oracle.c -- CPU restatement of the reference PBF path (see pbf_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the gfx950 library and the timed
 * CPU baseline; never part of the product.
 *
 * Build (oracle/Makefile):
 *   ORC_REAL=double  -> parity oracle (FP64 arithmetic, FP32 storage points)
 *   ORC_REAL=float   -> CPU baseline (FP32, OpenMP)
 * Both with -ffp-contract=off, so every FP32 expression is evaluated as
 * written (no FMA), and without -ffast-math.
 */
#include "pbf_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define ORC_STR2(x) #x
#define ORC_STR(x) ORC_STR2(x)

const char* orc_precision(void) { return ORC_STR(ORC_REAL); }

/* sph::pi of the reference is a float constant (include/pbf/sph_helpers.cuh:8). */
static const float kPiF = 3.14159265358979323846f;

static inline int valid_count(int c, int maxN) { return c >= 0 && c <= maxN; }

/* ------------------------------------------------------------------------ */
/* SPH helpers in the working precision                                      */
/* ------------------------------------------------------------------------ */

/* poly6FromDistanceSquared (sph_helpers.cuh:13-34): the normalisation is
 * divided by h one factor at a time. */
static inline real poly6_r(real d2, real h) {
    const real h2 = h * h;
    if (d2 >= h2) return (real)0;
    const real norm = (real)315 / ((real)64 * (real)kPiF) / h / h / h;
    const real rd = (real)1 - d2 / h2;
    return norm * (rd * rd * rd);
}

/* spikyGradientFromDistanceSquared (sph_helpers.cuh:48-81). */
static inline void spiky_r(real dx, real dy, real dz, real d2, real h, real g[3]) {
    const real r = (real)sqrt((double)d2);
    if (r <= (real)0 || r >= h) {
        g[0] = g[1] = g[2] = (real)0;
        return;
    }
    const real norm = (real)(-45) / (real)kPiF / h / h / h / h;
    const real rd = (real)1 - r / h;
    const real mag = norm * (rd * rd);
    g[0] = dx / r * mag;
    g[1] = dy / r * mag;
    g[2] = dz / r * mag;
}

/* computeArtificialPressure (include/pbf/pbf_solver_kernels.cuh:24-47). */
static inline real scorr_r(real d2, real h, real k, int n, real dq) {
    if (k == (real)0 || n <= 0) return (real)0;
    const real h2 = h * h;
    const real dq2 = dq * dq;
    if (d2 >= h2 || dq2 >= h2) return (real)0;
    const real ratio = (real)pow((double)((h2 - d2) / (h2 - dq2)), 3.0);
    return -k * (real)pow((double)ratio, (double)n);
}

float orc_poly6(orc_f3 d, float h) {
    const real dx = d.x, dy = d.y, dz = d.z;
    return (float)poly6_r(dx * dx + dy * dy + dz * dz, (real)h);
}

void orc_spiky(orc_f3 d, float h, orc_f3* out) {
    const real dx = d.x, dy = d.y, dz = d.z;
    real g[3];
    spiky_r(dx, dy, dz, dx * dx + dy * dy + dz * dz, (real)h, g);
    out->x = (float)g[0];
    out->y = (float)g[1];
    out->z = (float)g[2];
}

float orc_scorr(float d2, float h, float k, int n, float dq) {
    return (float)scorr_r((real)d2, (real)h, (real)k, n, (real)dq);
}

/* ------------------------------------------------------------------------ */
/* Grid                                                                      */
/* ------------------------------------------------------------------------ */

/* SpatialGrid::initialize validation and dimensions (spatial_grid.cu:10-64). */
int orc_grid_dims(size_t maxParticles, orc_f3 mn, orc_f3 mx, float cs, orc_i3* dims,
                  size_t* cells) {
    if (maxParticles == 0) return ORC_INVALID;
    if (!isfinite(cs) || cs <= 0.0f) return ORC_INVALID;
    if (!isfinite(mn.x) || !isfinite(mn.y) || !isfinite(mn.z) || !isfinite(mx.x) ||
        !isfinite(mx.y) || !isfinite(mx.z) || mx.x <= mn.x || mx.y <= mn.y || mx.z <= mn.z)
        return ORC_INVALID;
    if (maxParticles > (size_t)INT_MAX) return ORC_LENGTH;
    const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    int d[3];
    for (int a = 0; a < 3; ++a) {
        const double count = ceil(((double)hi[a] - (double)lo[a]) / (double)cs);
        if (!isfinite(count) || count < 1.0 || count > (double)INT_MAX) return ORC_LENGTH;
        d[a] = (int)count;
    }
    const size_t maxCells = (size_t)UINT32_MAX;
    const size_t cx = (size_t)d[0], cy = (size_t)d[1], cz = (size_t)d[2];
    if (cx > maxCells / cy || cx * cy > maxCells / cz) return ORC_LENGTH;
    dims->x = d[0];
    dims->y = d[1];
    dims->z = d[2];
    *cells = cx * cy * cz;
    return ORC_OK;
}

/* static_cast<int>(float) as CUDA performs it: saturating, NaN -> 0. */
static inline int sat_f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT_MAX;
    if (f < -2147483648.0f) return INT_MIN;
    return (int)f;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* positionToCell + clampCellToGrid (spatial_grid.cuh:65-102), FP32. */
static inline orc_i3 cell_of(orc_f4 p, orc_f3 mn, float cs, orc_i3 g) {
    orc_i3 c;
    const float qx = (p.x - mn.x) / cs;
    const float qy = (p.y - mn.y) / cs;
    const float qz = (p.z - mn.z) / cs;
    c.x = clampi(sat_f2i(floorf(qx)), 0, g.x - 1);
    c.y = clampi(sat_f2i(floorf(qy)), 0, g.y - 1);
    c.z = clampi(sat_f2i(floorf(qz)), 0, g.z - 1);
    return c;
}

/* cellToKey (spatial_grid.cuh:79-86): 64-bit arithmetic, truncated. */
static inline uint32_t key_of(int x, int y, int z, orc_i3 g) {
    return (uint32_t)((uint64_t)x + (uint64_t)y * (uint64_t)g.x +
                      (uint64_t)z * (uint64_t)g.x * (uint64_t)g.y);
}

/* computeParticleKeys (spatial_grid.cu:100-120). */
void orc_cell_keys(const orc_f4* p, size_t n, orc_f3 mn, float cs, orc_i3 g, uint32_t* keys) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const orc_i3 c = cell_of(p[i], mn, cs, g);
        keys[i] = key_of(c.x, c.y, c.z, g);
    }
}

/* Stable ascending sort of slots by key: a counting sort, which keeps equal
 * keys in slot order exactly like the reference's SortPairs over (key, slot)
 * (spatial_grid.cu:188-196, pinned by std::stable_sort in
 * test/test_spatial_grid.cu:104-142). */
static void stable_sort_by_key(const uint32_t* keys, size_t n, size_t cells, uint32_t* sortedKeys,
                               uint32_t* sortedIdx, uint32_t* counts /* cells + 1 */) {
    memset(counts, 0, (cells + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) counts[keys[i] + 1]++;
    for (size_t c = 0; c < cells; ++c) counts[c + 1] += counts[c];
    for (size_t i = 0; i < n; ++i) {
        const uint32_t pos = counts[keys[i]]++;
        sortedKeys[pos] = keys[i];
        sortedIdx[pos] = (uint32_t)i;
    }
}

/* buildCellRanges after the -1 fill (spatial_grid.cu:123-137, :203-213). */
static void cell_ranges(const uint32_t* sk, size_t n, size_t cells, int* cs, int* ce) {
    for (size_t c = 0; c < cells; ++c) cs[c] = ce[c] = -1;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t k = sk[i];
        if (i == 0 || sk[i - 1] != k) cs[k] = (int)i;
        if (i == n - 1 || sk[i + 1] != k) ce[k] = (int)(i + 1);
    }
}

void orc_grid_build(const orc_f4* p, size_t n, orc_f3 mn, float cs, orc_i3 g, size_t cells,
                    uint32_t* sortedKeys, uint32_t* sortedIdx, int* cellStart, int* cellEnd) {
    uint32_t* keys = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* counts = (uint32_t*)malloc((cells + 1) * sizeof(uint32_t));
    orc_cell_keys(p, n, mn, cs, g, keys);
    stable_sort_by_key(keys, n, cells, sortedKeys, sortedIdx, counts);
    cell_ranges(sortedKeys, n, cells, cellStart, cellEnd);
    free(keys);
    free(counts);
}

/* ------------------------------------------------------------------------ */
/* Neighbour search (src/pbf/neighbors.cu:3-91)                              */
/* ------------------------------------------------------------------------ */

/* Sweep order z, y, x over [cell - R, cell + R] clamped in double; inclusive
 * FP32 radius test (dx*dx + dy*dy) + dz*dz <= h*h; self excluded by index.
 * emit(i, count, j) stores the count-th neighbour when count < maxN. */
static int sweep_neighbors(const orc_f4* p, const int* cellStart, const int* cellEnd, orc_i3 g,
                           orc_f3 mn, float cs, size_t i, float h, uint32_t* out, int maxN,
                           size_t stride, size_t base) {
    const orc_f4 pi = p[i];
    const orc_i3 c = cell_of(pi, mn, cs, g);
    const float h2 = h * h;
    const double R = ceil((double)h / (double)cs);
    const int x0 = (int)fmax(0.0, c.x - R), y0 = (int)fmax(0.0, c.y - R),
              z0 = (int)fmax(0.0, c.z - R);
    const int x1 = (int)fmin((double)(g.x - 1), c.x + R),
              y1 = (int)fmin((double)(g.y - 1), c.y + R),
              z1 = (int)fmin((double)(g.z - 1), c.z + R);
    int count = 0;
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const uint32_t key = key_of(x, y, z, g);
                const int s = cellStart[key];
                if (s == -1) continue;
                const int e = cellEnd[key];
                for (int j = s; j < e; ++j) {
                    if ((size_t)j == i) continue;
                    const orc_f4 pj = p[j];
                    const float dx = pi.x - pj.x;
                    const float dy = pi.y - pj.y;
                    const float dz = pi.z - pj.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 <= h2) {
                        if (count < maxN) out[base + (size_t)count * stride] = (uint32_t)j;
                        ++count;
                    }
                }
            }
    return count;
}

int orc_find_neighbors(const orc_f4* p, const int* cellStart, const int* cellEnd, orc_i3 g,
                       orc_f3 mn, float cs, size_t n, float h, uint32_t* nbr, int* cnt, int maxN,
                       size_t stride) {
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
    for (long long i = 0; i < (long long)n; ++i) {
        const int c = sweep_neighbors(p, cellStart, cellEnd, g, mn, cs, (size_t)i, h, nbr, maxN,
                                      stride, (size_t)i);
        cnt[i] = c;
        if (c > maxN) overflow |= 1;
    }
    return overflow;
}

/* ------------------------------------------------------------------------ */
/* Particle kernels                                                          */
/* ------------------------------------------------------------------------ */

/* integrate (src/pbf/integration.cu:6-32). */
void orc_integrate(const orc_f4* x, orc_f4* p, orc_f4* v, size_t n, float dt, orc_f3 g) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const real vx = (real)v[i].x + (real)g.x * (real)dt;
        const real vy = (real)v[i].y + (real)g.y * (real)dt;
        const real vz = (real)v[i].z + (real)g.z * (real)dt;
        orc_f4 pi;
        pi.x = (float)((real)x[i].x + vx * (real)dt);
        pi.y = (float)((real)x[i].y + vy * (real)dt);
        pi.z = (float)((real)x[i].z + vz * (real)dt);
        pi.w = x[i].w;
        v[i].x = (float)vx;
        v[i].y = (float)vy;
        v[i].z = (float)vz;
        p[i] = pi;
    }
}

/* computeLambda (src/pbf/pbf_solver_kernels.cu:26-105). */
void orc_lambda(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN, size_t n,
                float h, float mass, float rho0, float eps, float* lambda, size_t stride) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        const int c = cnt[i];
        if (!valid_count(c, maxN)) {
            lambda[i] = NAN;
            continue;
        }
        const orc_f4 pi = p[i];
        const real gs = (real)mass / (real)rho0;
        real rho = (real)mass * poly6_r(0, (real)h);
        real gix = 0, giy = 0, giz = 0, sg = 0;
        for (int q = 0; q < c; ++q) {
            const orc_f4 pj = p[nbr[(size_t)q * stride + i]];
            const real dx = (real)pi.x - (real)pj.x;
            const real dy = (real)pi.y - (real)pj.y;
            const real dz = (real)pi.z - (real)pj.z;
            const real d2 = dx * dx + dy * dy + dz * dz;
            rho += (real)mass * poly6_r(d2, (real)h);
            real gw[3];
            spiky_r(dx, dy, dz, d2, (real)h, gw);
            const real gx = -gs * gw[0], gy = -gs * gw[1], gz = -gs * gw[2];
            sg += gx * gx + gy * gy + gz * gz;
            gix -= gx;
            giy -= gy;
            giz -= gz;
        }
        sg += gix * gix + giy * giy + giz * giz;
        const real C = rho / (real)rho0 - (real)1;
        lambda[i] = (float)(-C / (sg + (real)eps));
    }
}

/* computeDeltaPosition (src/pbf/pbf_solver_kernels.cu:108-182). */
void orc_delta(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN,
               const float* lambda, size_t n, float h, float mass, float rho0, float k, int sn,
               float dq, orc_f4* delta, size_t stride) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        const int c = cnt[i];
        if (!valid_count(c, maxN)) {
            delta[i].x = delta[i].y = delta[i].z = NAN;
            delta[i].w = 0.0f;
            continue;
        }
        const orc_f4 pi = p[i];
        const real li = (real)lambda[i];
        real ax = 0, ay = 0, az = 0;
        for (int q = 0; q < c; ++q) {
            const uint32_t j = nbr[(size_t)q * stride + i];
            const orc_f4 pj = p[j];
            const real dx = (real)pi.x - (real)pj.x;
            const real dy = (real)pi.y - (real)pj.y;
            const real dz = (real)pi.z - (real)pj.z;
            const real d2 = dx * dx + dy * dy + dz * dz;
            real gw[3];
            spiky_r(dx, dy, dz, d2, (real)h, gw);
            const real ls = li + (real)lambda[j] + scorr_r(d2, (real)h, (real)k, sn, (real)dq);
            ax += ls * gw[0];
            ay += ls * gw[1];
            az += ls * gw[2];
        }
        const real cs = (real)mass / (real)rho0;
        delta[i].x = (float)(cs * ax);
        delta[i].y = (float)(cs * ay);
        delta[i].z = (float)(cs * az);
        delta[i].w = 0.0f;
    }
}

/* applyDeltaPosition (src/pbf/pbf_solver_kernels.cu:184-193): FP32 adds. */
void orc_apply_delta(orc_f4* p, const orc_f4* d, size_t n) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        p[i].x = p[i].x + d[i].x;
        p[i].y = p[i].y + d[i].y;
        p[i].z = p[i].z + d[i].z;
    }
}

/* updateVelocityAndPosition (src/pbf/pbf_solver_kernels.cu:195-211). */
void orc_update_velocity(orc_f4* x, const orc_f4* p, orc_f4* v, size_t n, float invDt) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        v[i].x = (float)(((real)p[i].x - (real)x[i].x) * (real)invDt);
        v[i].y = (float)(((real)p[i].y - (real)x[i].y) * (real)invDt);
        v[i].z = (float)(((real)p[i].z - (real)x[i].z) * (real)invDt);
        x[i] = p[i];
    }
}

/* applyXsphViscosity (src/pbf/pbf_solver_kernels.cu:213-266). */
void orc_xsph(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN, const orc_f4* vin,
              orc_f4* vout, size_t n, float h, float visc, size_t stride) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        const orc_f4 vi = vin[i];
        const int c = cnt[i];
        if (!valid_count(c, maxN)) {
            vout[i].x = vout[i].y = vout[i].z = NAN;
            vout[i].w = vi.w;
            continue;
        }
        if (visc == 0.0f) {
            vout[i] = vi;
            continue;
        }
        const orc_f4 pi = p[i];
        real cx = 0, cy = 0, cz = 0;
        for (int q = 0; q < c; ++q) {
            const uint32_t j = nbr[(size_t)q * stride + i];
            const orc_f4 pj = p[j], vj = vin[j];
            const real dx = (real)pi.x - (real)pj.x;
            const real dy = (real)pi.y - (real)pj.y;
            const real dz = (real)pi.z - (real)pj.z;
            const real w = poly6_r(dx * dx + dy * dy + dz * dz, (real)h);
            cx += ((real)vj.x - (real)vi.x) * w;
            cy += ((real)vj.y - (real)vi.y) * w;
            cz += ((real)vj.z - (real)vi.z) * w;
        }
        vout[i].x = (float)((real)vi.x + (real)visc * cx);
        vout[i].y = (float)((real)vi.y + (real)visc * cy);
        vout[i].z = (float)((real)vi.z + (real)visc * cz);
        vout[i].w = vi.w;
    }
}

/* computeVorticity (src/pbf/pbf_solver_kernels.cu:268-315). */
void orc_vorticity(const orc_f4* p, const orc_f4* v, const uint32_t* nbr, const int* cnt,
                   int maxN, size_t n, float h, orc_f4* omega, size_t stride) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        const int c = cnt[i];
        if (!valid_count(c, maxN)) {
            omega[i].x = omega[i].y = omega[i].z = NAN;
            omega[i].w = 0.0f;
            continue;
        }
        const orc_f4 pi = p[i], vi = v[i];
        real ox = 0, oy = 0, oz = 0;
        for (int q = 0; q < c; ++q) {
            const uint32_t j = nbr[(size_t)q * stride + i];
            const orc_f4 pj = p[j], vj = v[j];
            const real dx = (real)pi.x - (real)pj.x;
            const real dy = (real)pi.y - (real)pj.y;
            const real dz = (real)pi.z - (real)pj.z;
            real g[3];
            spiky_r(dx, dy, dz, dx * dx + dy * dy + dz * dz, (real)h, g);
            const real ux = (real)vj.x - (real)vi.x;
            const real uy = (real)vj.y - (real)vi.y;
            const real uz = (real)vj.z - (real)vi.z;
            ox += g[1] * uz - g[2] * uy;
            oy += g[2] * ux - g[0] * uz;
            oz += g[0] * uy - g[1] * ux;
        }
        omega[i].x = (float)ox;
        omega[i].y = (float)oy;
        omega[i].z = (float)oz;
        omega[i].w = 0.0f;
    }
}

/* applyVorticityConfinement (src/pbf/pbf_solver_kernels.cu:317-385). */
void orc_confine(const orc_f4* p, const uint32_t* nbr, const int* cnt, int maxN,
                 const orc_f4* omega, const orc_f4* vin, orc_f4* vout, size_t n, float h,
                 float dt, float eps, size_t stride) {
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        const orc_f4 vi = vin[i];
        const int c = cnt[i];
        if (!valid_count(c, maxN)) {
            vout[i].x = vout[i].y = vout[i].z = NAN;
            vout[i].w = vi.w;
            continue;
        }
        const orc_f4 pi = p[i], wi = omega[i];
        const real wil = (real)sqrt((double)((real)wi.x * wi.x + (real)wi.y * wi.y + (real)wi.z * wi.z));
        real ex = 0, ey = 0, ez = 0;
        for (int q = 0; q < c; ++q) {
            const uint32_t j = nbr[(size_t)q * stride + i];
            const orc_f4 pj = p[j], wj = omega[j];
            const real wjl =
                (real)sqrt((double)((real)wj.x * wj.x + (real)wj.y * wj.y + (real)wj.z * wj.z));
            const real dx = (real)pi.x - (real)pj.x;
            const real dy = (real)pi.y - (real)pj.y;
            const real dz = (real)pi.z - (real)pj.z;
            real g[3];
            spiky_r(dx, dy, dz, dx * dx + dy * dy + dz * dz, (real)h, g);
            const real md = wjl - wil;
            ex += md * g[0];
            ey += md * g[1];
            ez += md * g[2];
        }
        const real el = (real)sqrt((double)(ex * ex + ey * ey + ez * ez));
        real nx = 0, ny = 0, nz = 0;
        if (el > (real)1.0e-6f) {
            nx = ex / el;
            ny = ey / el;
            nz = ez / el;
        }
        const real s = (real)eps;
        const real fx = s * (ny * wi.z - nz * wi.y);
        const real fy = s * (nz * wi.x - nx * wi.z);
        const real fz = s * (nx * wi.y - ny * wi.x);
        vout[i].x = (float)((real)vi.x + (real)dt * fx);
        vout[i].y = (float)((real)vi.y + (real)dt * fy);
        vout[i].z = (float)((real)vi.z + (real)dt * fz);
        vout[i].w = vi.w;
    }
}

/* ------------------------------------------------------------------------ */
/* Collisions, FP32 as written in src/pbf/collision_kernels.cu              */
/* ------------------------------------------------------------------------ */

/* projectionTolerance (collision_kernels.cu:9-14). */
static inline float projection_tolerance(orc_f4 p, orc_f3 o, float radius) {
    float scale = fmaxf(radius, fmaxf(fabsf(p.x), fabsf(o.x)));
    scale = fmaxf(scale, fmaxf(fabsf(p.y), fabsf(o.y)));
    scale = fmaxf(scale, fmaxf(fabsf(p.z), fabsf(o.z)));
    return 8.0f * FLT_EPSILON * scale;
}

/* solveContainerKernel body (collision_kernels.cu:56-81). */
static inline int project_container(orc_f4* q, const orc_container* c, float r) {
    orc_f4 p = *q;
    const orc_f4 o = p;
    const float minX = c->min.x + r, maxX = c->max.x - r;
    const float minY = c->min.y + r;
    const float minZ = c->min.z + r, maxZ = c->max.z - r;
    p.x = fminf(fmaxf(p.x, minX), maxX);
    p.y = fmaxf(p.y, minY);
    p.z = fminf(fmaxf(p.z, minZ), maxZ);
    *q = p;
    return p.x != o.x || p.y != o.y || p.z != o.z;
}

/* solveSpheresKernel loop body (collision_kernels.cu:96-120). */
static inline int project_sphere(orc_f4* q, const orc_sphere* s, float r) {
    orc_f4 p = *q;
    const float md = s->radius + r;
    const float ox = p.x - s->center.x, oy = p.y - s->center.y, oz = p.z - s->center.z;
    const float d2 = ox * ox + oy * oy + oz * oz;
    if (!(d2 < md * md)) return 0;
    if (md - sqrtf(d2) <= projection_tolerance(p, s->center, md)) return 0;
    if (d2 == 0.0f) {
        p.x = s->center.x + md;
    } else {
        const float scale = md / sqrtf(d2);
        p.x = s->center.x + ox * scale;
        p.y = s->center.y + oy * scale;
        p.z = s->center.z + oz * scale;
    }
    *q = p;
    return 1;
}

/* solveBoxesKernel loop body (collision_kernels.cu:140-186). */
static inline int project_box(orc_f4* q, const orc_box* b, float r) {
    orc_f4 p = *q;
    const float mnx = b->center.x - b->half.x, mny = b->center.y - b->half.y,
                mnz = b->center.z - b->half.z;
    const float mxx = b->center.x + b->half.x, mxy = b->center.y + b->half.y,
                mxz = b->center.z + b->half.z;
    orc_f3 cl;
    cl.x = fminf(fmaxf(p.x, mnx), mxx);
    cl.y = fminf(fmaxf(p.y, mny), mxy);
    cl.z = fminf(fmaxf(p.z, mnz), mxz);
    const float ox = p.x - cl.x, oy = p.y - cl.y, oz = p.z - cl.z;
    const float d2 = ox * ox + oy * oy + oz * oz;
    if (d2 > 0.0f && d2 < r * r) {
        if (r - sqrtf(d2) <= projection_tolerance(p, cl, r)) return 0;
        const float scale = r / sqrtf(d2);
        p.x = cl.x + ox * scale;
        p.y = cl.y + oy * scale;
        p.z = cl.z + oz * scale;
        *q = p;
        return 1;
    }
    if (d2 == 0.0f) {
        const float dx = fminf(p.x - mnx, mxx - p.x);
        const float dy = fminf(p.y - mny, mxy - p.y);
        const float dz = fminf(p.z - mnz, mxz - p.z);
        if (dx <= dy && dx <= dz)
            p.x = p.x - mnx < mxx - p.x ? mnx - r : mxx + r;
        else if (dy <= dz)
            p.y = p.y - mny < mxy - p.y ? mny - r : mxy + r;
        else
            p.z = p.z - mnz < mxz - p.z ? mnz - r : mxz + r;
        *q = p;
        return 1;
    }
    return 0;
}

/* solvePlanesKernel loop body (collision_kernels.cu:203-231). */
static inline int project_plane(orc_f4* q, const orc_plane* pl, float r) {
    orc_f4 p = *q;
    const float l2 =
        pl->normal.x * pl->normal.x + pl->normal.y * pl->normal.y + pl->normal.z * pl->normal.z;
    if (l2 == 0.0f) return 0;
    const float ns = 1.0f / sqrtf(l2);
    const float nx = pl->normal.x * ns, ny = pl->normal.y * ns, nz = pl->normal.z * ns;
    const float dist =
        (p.x - pl->point.x) * nx + (p.y - pl->point.y) * ny + (p.z - pl->point.z) * nz;
    if (!(dist < r)) return 0;
    if (r - dist <= projection_tolerance(p, pl->point, r)) return 0;
    const float corr = r - dist;
    p.x += nx * corr;
    p.y += ny * corr;
    p.z += nz * corr;
    *q = p;
    return 1;
}

void orc_solve_container(orc_f4* p, size_t n, const orc_container* c, float r, int* flag) {
    int marked = 0;
    for (size_t i = 0; i < n; ++i) marked |= project_container(&p[i], c, r);
    if (marked && flag) *flag = 1;
}

void orc_solve_spheres(orc_f4* p, size_t n, const orc_sphere* s, size_t ns, float r, int* flag) {
    int marked = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < ns; ++k) marked |= project_sphere(&p[i], &s[k], r);
    if (marked && flag) *flag = 1;
}

void orc_solve_boxes(orc_f4* p, size_t n, const orc_box* b, size_t nb, float r, int* flag) {
    int marked = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < nb; ++k) marked |= project_box(&p[i], &b[k], r);
    if (marked && flag) *flag = 1;
}

void orc_solve_planes(orc_f4* p, size_t n, const orc_plane* pl, size_t np, float r, int* flag) {
    int marked = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < np; ++k) marked |= project_plane(&p[i], &pl[k], r);
    if (marked && flag) *flag = 1;
}

/* CollisionSystem::solve (collision_system.cu:133-172): whole passes of
 * container, spheres, boxes, planes over every particle until a pass marks
 * nothing; non-convergence after 32 passes. */
int orc_collision_solve(orc_f4* p, size_t n, const orc_container* c, const orc_sphere* s,
                        size_t ns, const orc_box* b, size_t nb, const orc_plane* pl, size_t np,
                        float r) {
    for (int pass = 0; pass < 32; ++pass) {
        int marked = 0;
#pragma omp parallel for schedule(static) reduction(| : marked)
        for (long long i = 0; i < (long long)n; ++i) {
            int m = project_container(&p[i], c, r);
            for (size_t k = 0; k < ns; ++k) m |= project_sphere(&p[i], &s[k], r);
            for (size_t k = 0; k < nb; ++k) m |= project_box(&p[i], &b[k], r);
            for (size_t k = 0; k < np; ++k) m |= project_plane(&p[i], &pl[k], r);
            marked |= m;
        }
        if (!marked) return ORC_OK;
    }
    return ORC_NONCONVERGENCE;
}

/* resolveContactVelocity (collision_kernels.cu:21-43). */
static inline void contact_velocity(float nx, float ny, float nz, orc_f4 vin, float e,
                                    float ts, orc_f4* v) {
    const float vn = v->x * nx + v->y * ny + v->z * nz;
    const float vinn = vin.x * nx + vin.y * ny + vin.z * nz;
    const float target = vinn < 0.0f ? -e * vinn : 0.0f;
    if (vn > target || (vn == target && vinn >= 0.0f)) return;
    const float tx = v->x - vn * nx, ty = v->y - vn * ny, tz = v->z - vn * nz;
    v->x = target * nx + ts * tx;
    v->y = target * ny + ts * ty;
    v->z = target * nz + ts * tz;
}

/* resolveVelocitiesKernel body (collision_kernels.cu:236-371). */
static void resolve_one(orc_f4 p, orc_f4 vin, orc_f4* v, const orc_container* c,
                        const orc_sphere* s, size_t ns, const orc_box* b, size_t nb,
                        const orc_plane* pl, size_t np, float r, float e, float f) {
    const float ts = 1.0f - f;
    const float eps = 1e-4f;
    const float minX = c->min.x + r, maxX = c->max.x - r;
    const float minY = c->min.y + r;
    const float minZ = c->min.z + r, maxZ = c->max.z - r;
    if (p.x <= minX + eps)
        contact_velocity(1.0f, 0.0f, 0.0f, vin, e, ts, v);
    else if (p.x >= maxX - eps)
        contact_velocity(-1.0f, 0.0f, 0.0f, vin, e, ts, v);
    if (p.y <= minY + eps) contact_velocity(0.0f, 1.0f, 0.0f, vin, e, ts, v);
    if (p.z <= minZ + eps)
        contact_velocity(0.0f, 0.0f, 1.0f, vin, e, ts, v);
    else if (p.z >= maxZ - eps)
        contact_velocity(0.0f, 0.0f, -1.0f, vin, e, ts, v);
    for (size_t k = 0; k < ns; ++k) {
        const float md = s[k].radius + r;
        const float ox = p.x - s[k].center.x, oy = p.y - s[k].center.y, oz = p.z - s[k].center.z;
        const float d2 = ox * ox + oy * oy + oz * oz;
        const float cd = md + eps;
        if (d2 > 0.0f && d2 <= cd * cd) {
            const float sc = 1.0f / sqrtf(d2);
            contact_velocity(ox * sc, oy * sc, oz * sc, vin, e, ts, v);
        }
    }
    for (size_t k = 0; k < nb; ++k) {
        const orc_box* bx = &b[k];
        const float mnx = bx->center.x - bx->half.x, mny = bx->center.y - bx->half.y,
                    mnz = bx->center.z - bx->half.z;
        const float mxx = bx->center.x + bx->half.x, mxy = bx->center.y + bx->half.y,
                    mxz = bx->center.z + bx->half.z;
        const float cx = fminf(fmaxf(p.x, mnx), mxx);
        const float cy = fminf(fmaxf(p.y, mny), mxy);
        const float cz = fminf(fmaxf(p.z, mnz), mxz);
        const float ox = p.x - cx, oy = p.y - cy, oz = p.z - cz;
        const float d2 = ox * ox + oy * oy + oz * oz;
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        const float rc = r + eps;
        if (d2 > 0.0f && d2 <= rc * rc) {
            const float sc = 1.0f / sqrtf(d2);
            nx = ox * sc;
            ny = oy * sc;
            nz = oz * sc;
        } else if (d2 == 0.0f) {
            const float dx = fminf(p.x - mnx, mxx - p.x);
            const float dy = fminf(p.y - mny, mxy - p.y);
            const float dz = fminf(p.z - mnz, mxz - p.z);
            if (dx <= dy && dx <= dz)
                nx = p.x - mnx < mxx - p.x ? -1.0f : 1.0f;
            else if (dy <= dz)
                ny = p.y - mny < mxy - p.y ? -1.0f : 1.0f;
            else
                nz = p.z - mnz < mxz - p.z ? -1.0f : 1.0f;
        } else {
            continue;
        }
        contact_velocity(nx, ny, nz, vin, e, ts, v);
    }
    for (size_t k = 0; k < np; ++k) {
        const float l2 = pl[k].normal.x * pl[k].normal.x + pl[k].normal.y * pl[k].normal.y +
                         pl[k].normal.z * pl[k].normal.z;
        if (l2 == 0.0f) continue;
        const float ns2 = 1.0f / sqrtf(l2);
        const float nx = pl[k].normal.x * ns2, ny = pl[k].normal.y * ns2, nz = pl[k].normal.z * ns2;
        const float dist = (p.x - pl[k].point.x) * nx + (p.y - pl[k].point.y) * ny +
                           (p.z - pl[k].point.z) * nz;
        if (dist <= r + eps) contact_velocity(nx, ny, nz, vin, e, ts, v);
    }
}

void orc_resolve_velocities(const orc_f4* x, const orc_f4* vin, orc_f4* v, size_t n,
                            const orc_container* c, const orc_sphere* s, size_t ns,
                            const orc_box* b, size_t nb, const orc_plane* pl, size_t np, float r,
                            float e, float f) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i)
        resolve_one(x[i], vin[i], &v[i], c, s, ns, b, nb, pl, np, r, e, f);
}

/* ------------------------------------------------------------------------ */
/* Solver (src/pbf/pbf_solver.cu)                                            */
/* ------------------------------------------------------------------------ */

#define ORC_MAX_NEIGHBORS 256 /* kMaxNeighbors, pbf_solver.cu:12 */

struct orc_solver {
    size_t maxN, count, stride;
    orc_params P;
    orc_f3 mn, mx;
    orc_i3 dims;
    size_t cells;
    orc_container container;
    orc_sphere* spheres;
    size_t ns;
    orc_box* boxes;
    size_t nb;
    orc_plane* planes;
    size_t np;
    orc_f4 *x, *p, *v, *vin, *xs, *omega, *delta, *tmp4;
    uint32_t *id, *keys, *sk, *sidx, *counts, *nbr, *tmpu;
    int *cs, *ce, *cnt;
    float* lambda;
};

static int params_valid(const orc_params* p) {
    if (!isfinite(p->dt) || p->dt <= 0.0f || !isfinite(p->restDensity) || p->restDensity <= 0.0f ||
        !isfinite(p->particleMass) || p->particleMass <= 0.0f || !isfinite(p->particleRadius) ||
        p->particleRadius <= 0.0f || !isfinite(p->collisionRestitution) ||
        p->collisionRestitution < 0.0f || p->collisionRestitution > 1.0f ||
        !isfinite(p->collisionFriction) || p->collisionFriction < 0.0f ||
        p->collisionFriction > 1.0f || !isfinite(p->smoothingRadius) ||
        p->smoothingRadius <= 0.0f || !isfinite(p->lambdaEpsilon) || p->lambdaEpsilon <= 0.0f ||
        !isfinite(p->scorrK) || p->scorrK < 0.0f ||
        (p->scorrK > 0.0f && (p->scorrN <= 0 || !isfinite(p->scorrDeltaQ) ||
                              p->scorrDeltaQ < 0.0f || p->scorrDeltaQ >= p->smoothingRadius)) ||
        !isfinite(p->xsphViscosity) || p->xsphViscosity < 0.0f ||
        !isfinite(p->vorticityStrength) || p->vorticityStrength < 0.0f ||
        !isfinite(p->gravity.x) || !isfinite(p->gravity.y) || !isfinite(p->gravity.z) ||
        p->solverIterations <= 0 || p->substeps <= 0)
        return 0;
    const float sdt = p->dt / (float)p->substeps;
    if (!isfinite(sdt) || sdt <= 0.0f || !isfinite(1.0f / sdt)) return 0;
    return 1;
}

void orc_solver_destroy(orc_solver* s) {
    if (!s) return;
    free(s->spheres);
    free(s->boxes);
    free(s->planes);
    free(s->x); free(s->p); free(s->v); free(s->vin); free(s->xs); free(s->omega);
    free(s->delta); free(s->tmp4);
    free(s->id); free(s->keys); free(s->sk); free(s->sidx); free(s->counts); free(s->nbr);
    free(s->tmpu);
    free(s->cs); free(s->ce); free(s->cnt); free(s->lambda);
    free(s);
}

int orc_solver_create(orc_solver** out, size_t maxN, orc_f3 mn, orc_f3 mx,
                      const orc_params* params) {
    *out = NULL;
    if (maxN == 0) return ORC_INVALID;
    if (maxN > (size_t)INT_MAX) return ORC_LENGTH;
    if (!params_valid(params)) return ORC_INVALID;
    const double width = (double)mx.x - (double)mn.x, depth = (double)mx.z - (double)mn.z;
    const double diameter = 2.0 * (double)params->particleRadius;
    if (width < diameter || depth < diameter) return ORC_INVALID;
    orc_i3 dims;
    size_t cells;
    const int st = orc_grid_dims(maxN, mn, mx, params->smoothingRadius, &dims, &cells);
    if (st != ORC_OK) return st;
    orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
    s->maxN = maxN;
    s->stride = ((maxN + 255) / 256) * 256;
    s->P = *params;
    s->mn = mn;
    s->mx = mx;
    s->dims = dims;
    s->cells = cells;
    s->container.min = mn;
    s->container.max = mx;
    const size_t f4 = maxN * sizeof(orc_f4);
    s->x = (orc_f4*)malloc(f4); s->p = (orc_f4*)malloc(f4); s->v = (orc_f4*)malloc(f4);
    s->vin = (orc_f4*)malloc(f4); s->xs = (orc_f4*)malloc(f4); s->omega = (orc_f4*)malloc(f4);
    s->delta = (orc_f4*)malloc(f4); s->tmp4 = (orc_f4*)malloc(f4 * 4);
    s->id = (uint32_t*)malloc(maxN * 4); s->keys = (uint32_t*)malloc(maxN * 4);
    s->sk = (uint32_t*)malloc(maxN * 4); s->sidx = (uint32_t*)malloc(maxN * 4);
    s->tmpu = (uint32_t*)malloc(maxN * 4);
    s->counts = (uint32_t*)malloc((cells + 1) * 4);
    s->nbr = (uint32_t*)malloc(s->stride * ORC_MAX_NEIGHBORS * 4);
    s->cs = (int*)malloc(cells * sizeof(int)); s->ce = (int*)malloc(cells * sizeof(int));
    s->cnt = (int*)malloc(maxN * sizeof(int)); s->lambda = (float*)malloc(maxN * sizeof(float));
    if (!s->x || !s->nbr || !s->tmp4 || !s->counts || !s->cs || !s->ce) {
        orc_solver_destroy(s);
        return ORC_LENGTH;
    }
    *out = s;
    return ORC_OK;
}

int orc_solver_set_particles(orc_solver* s, const orc_f4* pos, const orc_f4* vel, size_t n) {
    if (n > s->maxN) return ORC_OUT_OF_RANGE;
    if (n == 0) {
        s->count = 0;
        return ORC_OK;
    }
    if (!pos || !vel) return ORC_INVALID;
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(pos[i].x) || !isfinite(pos[i].y) || !isfinite(pos[i].z) ||
            !isfinite(vel[i].x) || !isfinite(vel[i].y) || !isfinite(vel[i].z))
            return ORC_INVALID;
    memcpy(s->x, pos, n * sizeof(orc_f4));
    memcpy(s->v, vel, n * sizeof(orc_f4));
    for (size_t i = 0; i < n; ++i) s->id[i] = (uint32_t)i;
    s->count = n;
    return ORC_OK;
}

static void* dup_array(const void* src, size_t bytes) {
    if (bytes == 0) return NULL;
    void* d = malloc(bytes);
    memcpy(d, src, bytes);
    return d;
}

int orc_solver_set_colliders(orc_solver* s, const orc_sphere* sp, size_t ns, const orc_box* b,
                             size_t nb, const orc_plane* pl, size_t np) {
    free(s->spheres);
    free(s->boxes);
    free(s->planes);
    s->spheres = (orc_sphere*)dup_array(sp, ns * sizeof(orc_sphere));
    s->boxes = (orc_box*)dup_array(b, nb * sizeof(orc_box));
    s->planes = (orc_plane*)dup_array(pl, np * sizeof(orc_plane));
    s->ns = ns;
    s->nb = nb;
    s->np = np;
    return ORC_OK;
}

/* SpatialGrid::build + reorderWorkingSet (pbf_solver.cu:154-172). */
static void build_and_reorder(orc_solver* s) {
    const size_t n = s->count;
    orc_cell_keys(s->p, n, s->mn, s->P.smoothingRadius, s->dims, s->keys);
    stable_sort_by_key(s->keys, n, s->cells, s->sk, s->sidx, s->counts);
    cell_ranges(s->sk, n, s->cells, s->cs, s->ce);
    orc_f4* arrays[4] = {s->x, s->p, s->v, s->vin};
    for (int a = 0; a < 4; ++a) {
        orc_f4* src = arrays[a];
        orc_f4* dst = s->tmp4 + (size_t)a * s->maxN;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) dst[i] = src[s->sidx[i]];
        memcpy(src, dst, n * sizeof(orc_f4));
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) s->tmpu[i] = s->id[s->sidx[i]];
    memcpy(s->id, s->tmpu, n * sizeof(uint32_t));
}

static int neighbors(orc_solver* s) {
    const int overflow =
        orc_find_neighbors(s->p, s->cs, s->ce, s->dims, s->mn, s->P.smoothingRadius, s->count,
                           s->P.smoothingRadius, s->nbr, s->cnt, ORC_MAX_NEIGHBORS, s->stride);
    return overflow ? ORC_OVERFLOW : ORC_OK;
}

/* PBFSolver::step (pbf_solver.cu:186-328). */
int orc_solver_step(orc_solver* s) {
    const size_t n = s->count;
    if (n == 0) return ORC_OK;
    const orc_params* P = &s->P;
    const float sdt = P->dt / (float)P->substeps;
    const float invDt = 1.0f / sdt;
    const float h = P->smoothingRadius;
    const size_t S = s->stride;
    const int M = ORC_MAX_NEIGHBORS;
    for (int sub = 0; sub < P->substeps; ++sub) {
        orc_integrate(s->x, s->p, s->v, n, sdt, P->gravity);
        memcpy(s->vin, s->v, n * sizeof(orc_f4));
        for (int it = 0; it < P->solverIterations; ++it) {
            build_and_reorder(s);
            int st = neighbors(s);
            if (st != ORC_OK) return st;
            orc_lambda(s->p, s->nbr, s->cnt, M, n, h, P->particleMass, P->restDensity,
                       P->lambdaEpsilon, s->lambda, S);
            orc_delta(s->p, s->nbr, s->cnt, M, s->lambda, n, h, P->particleMass, P->restDensity,
                      P->scorrK, P->scorrN, P->scorrDeltaQ, s->delta, S);
            orc_apply_delta(s->p, s->delta, n);
            st = orc_collision_solve(s->p, n, &s->container, s->spheres, s->ns, s->boxes, s->nb,
                                     s->planes, s->np, P->particleRadius);
            if (st != ORC_OK) return st;
        }
        if (P->xsphViscosity != 0.0f || P->vorticityStrength != 0.0f) {
            build_and_reorder(s);
            const int st = neighbors(s);
            if (st != ORC_OK) return st;
        }
        orc_update_velocity(s->x, s->p, s->v, n, invDt);
        if (P->xsphViscosity != 0.0f) {
            orc_xsph(s->p, s->nbr, s->cnt, M, s->v, s->xs, n, h, P->xsphViscosity, S);
            orc_f4* t = s->v; s->v = s->xs; s->xs = t;
        }
        if (P->vorticityStrength != 0.0f) {
            orc_vorticity(s->p, s->v, s->nbr, s->cnt, M, n, h, s->omega, S);
            orc_confine(s->p, s->nbr, s->cnt, M, s->omega, s->v, s->xs, n, h, sdt,
                        P->vorticityStrength, S);
            orc_f4* t = s->v; s->v = s->xs; s->xs = t;
        }
        orc_resolve_velocities(s->x, s->vin, s->v, n, &s->container, s->spheres, s->ns, s->boxes,
                               s->nb, s->planes, s->np, P->particleRadius,
                               P->collisionRestitution, P->collisionFriction);
    }
    return ORC_OK;
}

int orc_solver_run(orc_solver* s, size_t frames) {
    for (size_t f = 0; f < frames; ++f) {
        const int st = orc_solver_step(s);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

int orc_solver_get(const orc_solver* s, orc_f4* pos, orc_f4* vel, size_t n) {
    if (n != s->count) return ORC_OUT_OF_RANGE;
    for (size_t i = 0; i < n; ++i) {
        if (pos) pos[s->id[i]] = s->x[i];
        if (vel) vel[s->id[i]] = s->v[i];
    }
    return ORC_OK;
}

int orc_solver_working_set(const orc_solver* s, orc_f4* predicted, uint32_t* ids, size_t n) {
    if (n != s->count) return ORC_OUT_OF_RANGE;
    if (predicted) memcpy(predicted, s->p, n * sizeof(orc_f4));
    if (ids) memcpy(ids, s->id, n * sizeof(uint32_t));
    return ORC_OK;
}
