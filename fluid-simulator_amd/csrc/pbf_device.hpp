// Device-side building blocks shared by every gfx950 kernel of the library.
//
// Numerics follow the reference formula by formula; the reference sites are
// cited per function.  Two rules keep results reproducible:
//  * index decisions (cell hash, neighbour radius test) and the collision
//    projections are evaluated with FP32 multiplies and adds that are never
//    contracted into FMAs, so they are bit-identical to a host evaluation of
//    the same expression (the oracle does exactly that);
//  * the SPH sums use FMAs, spelled out wherever an expression holds two
//    products (dist2_fma, cross_fma) so that every inlined copy of a pass
//    rounds alike; they are compared to the FP64 oracle within a tolerance.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>

#include "pbf_amd.h"
#include "pbf_colliders.hpp"

namespace pbfdev {

constexpr float kPi = 3.14159265358979323846f;
constexpr int kMaxCollisionPasses = 32;  // collision_system.cu:148

__device__ __forceinline__ float4 ld4(const pbf_float4* p, size_t i) {
    return reinterpret_cast<const float4*>(p)[i];
}
__device__ __forceinline__ void st4(pbf_float4* p, size_t i, float4 v) {
    reinterpret_cast<float4*>(p)[i] = v;
}

__device__ __forceinline__ float qnan() { return __builtin_nanf(""); }

// hasValidNeighborCount (src/pbf/pbf_solver_kernels.cu:8-14), without the
// device assert: an invalid count NaN-poisons the outputs instead.
__device__ __forceinline__ bool valid_count(int c, int maxNeighbors) {
    return c >= 0 && c <= maxNeighbors;
}

// ---------------------------------------------------------------------------
// Cell hash (include/pbf/spatial_grid.cuh:65-102, src/pbf/spatial_grid.cu:100-120)
// ---------------------------------------------------------------------------

// Float -> int32 with saturation and NaN -> 0: the conversion CUDA's
// static_cast<int>(float) performs, which the +-1e30 grid test pins
// (test/test_spatial_grid.cu:682-689).  Written out so the behaviour does
// not rest on an out-of-range conversion.
__device__ __forceinline__ int sat_f2i(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return INT_MAX;
    if (f < -2147483648.0f) return INT_MIN;
    return static_cast<int>(f);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// positionToCell + clampCellToGrid: floorf((p - min) / cellSize) with an
// IEEE division, saturated, then clamped into the grid.
__device__ __forceinline__ int3 cell_of(float4 p, pbf_float3 mn, float cs, pbf_int3 g) {
#pragma clang fp contract(off)
    int3 c;
    c.x = clampi(sat_f2i(floorf((p.x - mn.x) / cs)), 0, g.x - 1);
    c.y = clampi(sat_f2i(floorf((p.y - mn.y) / cs)), 0, g.y - 1);
    c.z = clampi(sat_f2i(floorf((p.z - mn.z) / cs)), 0, g.z - 1);
    return c;
}

// cellToKey: x + y*gx + z*gx*gy in 64-bit, truncated to 32 bits.
__device__ __forceinline__ uint32_t key_of(int x, int y, int z, pbf_int3 g) {
    const uint64_t gx = static_cast<uint64_t>(g.x);
    const uint64_t gy = static_cast<uint64_t>(g.y);
    return static_cast<uint32_t>(static_cast<uint64_t>(x) + static_cast<uint64_t>(y) * gx +
                                 static_cast<uint64_t>(z) * gx * gy);
}

__device__ __forceinline__ uint32_t hash_key(float4 p, pbf_float3 mn, float cs, pbf_int3 g) {
    const int3 c = cell_of(p, mn, cs, g);
    return key_of(c.x, c.y, c.z, g);
}

// Squared distance exactly as the reference writes it, (dx*dx + dy*dy) + dz*dz,
// with no FMA: the inclusive neighbour test d2 <= h*h (src/pbf/neighbors.cu:76)
// depends on its last bit (test/test_neighbors.cu:215-231).
__device__ __forceinline__ float dist2_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// The squared distance of the SPH sums, with its contraction spelled out:
// fma(dz, dz, fma(dx, dx, dy * dy)).  Written as dx*dx + dy*dy + dz*dz, the
// compiler may contract the first sum around either product, and it chooses
// per inlined copy: two forms of one pass (batched and single-scene, the
// list build's XSPH and the separate pass) would then round differently,
// and any change elsewhere in a kernel could move its results by an ulp.
__device__ __forceinline__ float dist2_fma(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
}
__device__ __forceinline__ double dist2_fma(double dx, double dy, double dz) {
    return __builtin_fma(dz, dz, __builtin_fma(dx, dx, dy * dy));
}
// a * b - c * d with its contraction spelled out (a component of a cross
// product), for the same reason: fma(a, b, -(c * d)).
__device__ __forceinline__ float cross_fma(float a, float b, float c, float d) {
    return __builtin_fmaf(a, b, -(c * d));
}

// (p* - x) / dt as the reference's velocity update stores it
// (pbf_solver_kernels.cu:207): rounded, never fused into the XSPH sums that
// read it back.
__device__ __forceinline__ float stored_velocity(float pc, float xc, float inv_dt) {
#pragma clang fp contract(off)
    return (pc - xc) * inv_dt;
}

// ---------------------------------------------------------------------------
// SPH kernels (include/pbf/sph_helpers.cuh:13-95)
// ---------------------------------------------------------------------------
struct Sph {
    float h;
    float h2;
    float inv_h;
    float inv_h2;
    float poly_norm;   // 315 / (64 pi) / h / h / h, divided step by step
    float spiky_norm;  // -45 / pi / h / h / h / h, divided step by step
};

__device__ __forceinline__ Sph sph_consts(float h) {
    Sph s;
    s.h = h;
    s.h2 = h * h;
    s.inv_h = 1.0f / h;
    s.inv_h2 = 1.0f / s.h2;
    // Dividing the normalisation by h one factor at a time keeps h = 1e-5 .. 1e5
    // representable (test/test_sph_helpers.cu:153-171).
    s.poly_norm = 315.0f / (64.0f * kPi) / h / h / h;
    s.spiky_norm = -45.0f / kPi / h / h / h / h;
    return s;
}

// poly6FromDistanceSquared: zero for d2 >= h^2, else norm * (1 - d2/h^2)^3.
// (h^2 - d2) / h^2 is used for 1 - d2/h^2: equal in exact arithmetic, never
// negative in FP32, and free of a per-neighbour division.
__device__ __forceinline__ float poly6_d2(float d2, const Sph& s) {
    if (d2 >= s.h2) return 0.0f;
    const float rd = (s.h2 - d2) * s.inv_h2;
    return s.poly_norm * (rd * rd * rd);
}

// One neighbour's XSPH term (applyXsphViscosity, pbf_solver_kernels.cu:
// 242-257): c += (v_j - v_i) W(p*_i - p*_j).  The list build's XSPH and the
// separate pass both call it, and agree bit for bit.
__device__ __forceinline__ void xsph_term(float3 pi, float3 vi, float3 pj, float3 vj, const Sph& s,
                                          float3& c) {
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float d2 = dist2_fma(dx, dy, dz);
    const float w = poly6_d2(d2, s);
    c.x += (vj.x - vi.x) * w;
    c.y += (vj.y - vi.y) * w;
    c.z += (vj.z - vi.z) * w;
}

// spikyGradientFromDistanceSquared: zero for r <= 0 or r >= h with
// r = sqrtf(d2) correctly rounded (the cut-off test of the reference), else
// norm * (1 - r/h)^2 * d / r.
__device__ __forceinline__ float3 spiky_grad(float dx, float dy, float dz, float d2, const Sph& s) {
    const float r = __builtin_sqrtf(d2);
    if (r <= 0.0f || r >= s.h) return make_float3(0.0f, 0.0f, 0.0f);
    const float rd = (s.h - r) * s.inv_h;
    const float scale = (s.spiky_norm * (rd * rd)) / r;
    return make_float3(dx * scale, dy * scale, dz * scale);
}

// spiky_grad with the hardware reciprocal square root (v_rsq_f32, 1 ulp) in
// place of the correctly rounded sqrt and division: about 25 fewer
// instructions per neighbour, a few ulp per term.  The step's neighbour
// passes use it; their sums are held to the FP64 oracle within 1e-5.  The
// cut-off r >= h is evaluated on the approximate r, which can only differ
// from the exact decision where (1 - r/h)^2 is below 1e-13.  Subnormal d2
// (particles closer than 1e-19) takes the exact path.
__device__ __forceinline__ float3 spiky_grad_rsq(float dx, float dy, float dz, float d2,
                                                 const Sph& s) {
    if (d2 < FLT_MIN) return spiky_grad(dx, dy, dz, d2, s);
    const float inv_r = __builtin_amdgcn_rsqf(d2);
    const float r = d2 * inv_r;
    if (r >= s.h) return make_float3(0.0f, 0.0f, 0.0f);
    const float rd = (s.h - r) * s.inv_h;
    const float scale = s.spiky_norm * (rd * rd) * inv_r;
    return make_float3(dx * scale, dy * scale, dz * scale);
}

// computeArtificialPressure (include/pbf/pbf_solver_kernels.cuh:24-47):
// -k * (((h^2 - d2) / (h^2 - dq^2))^3)^n, zero when disabled or out of range.
// The integer powers are repeated products, not powf.
struct Scorr {
    float k;       // 0 disables
    int n;
    float inv_den; // 1 / (h^2 - dq^2)
    bool enabled;
};

__device__ __forceinline__ Scorr scorr_consts(float h, float k, int n, float dq) {
    Scorr c;
    const float h2 = h * h;
    const float dq2 = dq * dq;
    c.enabled = !(k == 0.0f || n <= 0) && !(dq2 >= h2);
    c.k = k;
    c.n = n;
    c.inv_den = c.enabled ? 1.0f / (h2 - dq2) : 0.0f;
    return c;
}

__device__ __forceinline__ float ipow(float x, int n) {
    float r = 1.0f;
    float b = x;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ float scorr_value(float d2, float h2, const Scorr& c) {
    if (!c.enabled || d2 >= h2) return 0.0f;
    const float q = (h2 - d2) * c.inv_den;
    const float ratio = q * q * q;
    if (c.n == 4) {
        const float r2 = ratio * ratio;
        return -c.k * (r2 * r2);
    }
    return -c.k * ipow(ratio, c.n);
}

// ---------------------------------------------------------------------------
// Collision projections (src/pbf/collision_kernels.cu:7-234), FP32, no FMA.
// Each returns true when the reference would markCorrection().
// ---------------------------------------------------------------------------

// projectionTolerance (collision_kernels.cu:9-14).
__device__ __forceinline__ float projection_tolerance(float px, float py, float pz, float ox,
                                                      float oy, float oz, float radius) {
    float scale = fmaxf(radius, fmaxf(fabsf(px), fabsf(ox)));
    scale = fmaxf(scale, fmaxf(fabsf(py), fabsf(oy)));
    scale = fmaxf(scale, fmaxf(fabsf(pz), fabsf(oz)));
    return 8.0f * FLT_EPSILON * scale;
}

// solveContainerKernel body (collision_kernels.cu:47-82): 5 walls, open top.
__device__ __forceinline__ bool project_container(float4& p, const pbf_container& c, float r) {
#pragma clang fp contract(off)
    const float minX = c.min.x + r;
    const float maxX = c.max.x - r;
    const float minY = c.min.y + r;
    const float minZ = c.min.z + r;
    const float maxZ = c.max.z - r;
    const float ox = p.x, oy = p.y, oz = p.z;
    p.x = fminf(fmaxf(p.x, minX), maxX);
    p.y = fmaxf(p.y, minY);
    p.z = fminf(fmaxf(p.z, minZ), maxZ);
    return p.x != ox || p.y != oy || p.z != oz;
}

// solveSpheresKernel body for one sphere (collision_kernels.cu:96-120).
__device__ __forceinline__ bool project_sphere(float4& p, const pbf_sphere& s, float r) {
#pragma clang fp contract(off)
    const float minDistance = s.radius + r;
    const float ox = p.x - s.center.x;
    const float oy = p.y - s.center.y;
    const float oz = p.z - s.center.z;
    const float d2 = ox * ox + oy * oy + oz * oz;
    if (!(d2 < minDistance * minDistance)) return false;
    if (minDistance - __builtin_sqrtf(d2) <=
        projection_tolerance(p.x, p.y, p.z, s.center.x, s.center.y, s.center.z, minDistance))
        return false;
    if (d2 == 0.0f) {
        p.x = s.center.x + minDistance;
    } else {
        const float scale = minDistance / __builtin_sqrtf(d2);
        p.x = s.center.x + ox * scale;
        p.y = s.center.y + oy * scale;
        p.z = s.center.z + oz * scale;
    }
    return true;
}

// solveBoxesKernel body for one box (collision_kernels.cu:140-186): rounded
// box; an interior point goes to the nearest face, ties to +x first.
__device__ __forceinline__ bool project_box(float4& p, const pbf_box& b, float r) {
#pragma clang fp contract(off)
    const float mnx = b.center.x - b.halfExtents.x;
    const float mny = b.center.y - b.halfExtents.y;
    const float mnz = b.center.z - b.halfExtents.z;
    const float mxx = b.center.x + b.halfExtents.x;
    const float mxy = b.center.y + b.halfExtents.y;
    const float mxz = b.center.z + b.halfExtents.z;
    const float cx = fminf(fmaxf(p.x, mnx), mxx);
    const float cy = fminf(fmaxf(p.y, mny), mxy);
    const float cz = fminf(fmaxf(p.z, mnz), mxz);
    const float ox = p.x - cx;
    const float oy = p.y - cy;
    const float oz = p.z - cz;
    const float d2 = ox * ox + oy * oy + oz * oz;
    if (d2 > 0.0f && d2 < r * r) {
        if (r - __builtin_sqrtf(d2) <= projection_tolerance(p.x, p.y, p.z, cx, cy, cz, r))
            return false;
        const float scale = r / __builtin_sqrtf(d2);
        p.x = cx + ox * scale;
        p.y = cy + oy * scale;
        p.z = cz + oz * scale;
        return true;
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
        return true;
    }
    return false;
}

// solvePlanesKernel body for one plane (collision_kernels.cu:203-231).
__device__ __forceinline__ bool project_plane(float4& p, const pbf_plane& pl, float r) {
#pragma clang fp contract(off)
    const float l2 = pl.normal.x * pl.normal.x + pl.normal.y * pl.normal.y +
                     pl.normal.z * pl.normal.z;
    if (l2 == 0.0f) return false;
    const float ns = 1.0f / __builtin_sqrtf(l2);
    const float nx = pl.normal.x * ns;
    const float ny = pl.normal.y * ns;
    const float nz = pl.normal.z * ns;
    const float dist = (p.x - pl.point.x) * nx + (p.y - pl.point.y) * ny + (p.z - pl.point.z) * nz;
    if (!(dist < r)) return false;
    if (r - dist <= projection_tolerance(p.x, p.y, p.z, pl.point.x, pl.point.y, pl.point.z, r))
        return false;
    const float corr = r - dist;
    p.x += nx * corr;
    p.y += ny * corr;
    p.z += nz * corr;
    return true;
}

// The container projection of p's coordinates (project_container's
// arithmetic).  Skin lists are built around these points: the projection is
// a per-axis clamp, so it never moves two particles farther apart, and a
// pair within h before it is within h after it.
__device__ __forceinline__ float3 container_point(float4 p, const pbf_container& c, float r) {
    project_container(p, c, r);
    return make_float3(p.x, p.y, p.z);
}

// One pass of CollisionSystem::solve for one particle (collision_system.cu:150-169):
// container, then spheres, boxes, planes, each kernel in array order.
__device__ __forceinline__ bool collision_pass(float4& p, const Colliders& c, float r) {
    bool marked = project_container(p, *c.container, r);
    for (int i = 0; i < c.sphereCount; ++i) marked |= project_sphere(p, c.spheres[i], r);
    for (int i = 0; i < c.boxCount; ++i) marked |= project_box(p, c.boxes[i], r);
    for (int i = 0; i < c.planeCount; ++i) marked |= project_plane(p, c.planes[i], r);
    return marked;
}

// The reference repeats full passes until one pass marks nothing, at most 32.
// A pass is a pure per-particle map, and a particle it does not mark is left
// unchanged, so iterating each particle on its own to its first unmarked
// pass gives the same positions.  Returns false when all 32 passes marked
// (the reference throws std::runtime_error, collision_system.cu:171).
__device__ __forceinline__ bool collision_fixed_point(float4& p, const Colliders& c, float r) {
    for (int pass = 0; pass < kMaxCollisionPasses; ++pass) {
        if (!collision_pass(p, c, r)) return true;
    }
    return false;
}

// resolveContactVelocity (collision_kernels.cu:21-43).
__device__ __forceinline__ void contact_velocity(float nx, float ny, float nz, const float4& vin,
                                                 float restitution, float tangentScale,
                                                 float4& v) {
#pragma clang fp contract(off)
    const float vn = v.x * nx + v.y * ny + v.z * nz;
    const float vinn = vin.x * nx + vin.y * ny + vin.z * nz;
    const float target = vinn < 0.0f ? -restitution * vinn : 0.0f;
    if (vn > target || (vn == target && vinn >= 0.0f)) return;
    const float tx = v.x - vn * nx;
    const float ty = v.y - vn * ny;
    const float tz = v.z - vn * nz;
    v.x = target * nx + tangentScale * tx;
    v.y = target * ny + tangentScale * ty;
    v.z = target * nz + tangentScale * tz;
}

// The geometry of one contact of resolveVelocitiesKernel: its normal, and
// for normals that follow the position (spheres, and rounded boxes outside
// their core) the reciprocal distance 1 / |o| and the axes along which o
// follows the position (o = p - clamp(p) for boxes).
struct ContactGeom {
    float nx, ny, nz;
    bool radial;
    float inv_len;
    bool outX, outY, outZ;
};

// Calls visit(geom) for every contact of resolveVelocitiesKernel
// (collision_kernels.cu:236-371) that p touches, in the reference's order:
// the container walls (floor, -x or +x, -z or +z; the top is open), then
// every sphere, box and plane in array order.
template <class Visit>
__device__ __forceinline__ void for_each_contact(const float4& p, const Colliders& c, float r,
                                                 Visit visit) {
#pragma clang fp contract(off)
    constexpr float eps = 1e-4f;  // collisionEpsilon (collision_kernels.cu:254)
    const pbf_container& k = *c.container;
    const float minX = k.min.x + r;
    const float maxX = k.max.x - r;
    const float minY = k.min.y + r;
    const float minZ = k.min.z + r;
    const float maxZ = k.max.z - r;
    const auto wall = [&](float nx, float ny, float nz) {
        visit(ContactGeom{nx, ny, nz, false, 0.0f, false, false, false});
    };
    if (p.x <= minX + eps)
        wall(1.0f, 0.0f, 0.0f);
    else if (p.x >= maxX - eps)
        wall(-1.0f, 0.0f, 0.0f);
    if (p.y <= minY + eps) wall(0.0f, 1.0f, 0.0f);
    if (p.z <= minZ + eps)
        wall(0.0f, 0.0f, 1.0f);
    else if (p.z >= maxZ - eps)
        wall(0.0f, 0.0f, -1.0f);

    for (int i = 0; i < c.sphereCount; ++i) {
        const pbf_sphere s = c.spheres[i];
        const float md = s.radius + r;
        const float ox = p.x - s.center.x;
        const float oy = p.y - s.center.y;
        const float oz = p.z - s.center.z;
        const float d2 = ox * ox + oy * oy + oz * oz;
        const float cd = md + eps;
        if (d2 > 0.0f && d2 <= cd * cd) {
            const float sc = 1.0f / __builtin_sqrtf(d2);
            visit(ContactGeom{ox * sc, oy * sc, oz * sc, true, sc, true, true, true});
        }
    }
    for (int i = 0; i < c.boxCount; ++i) {
        const pbf_box b = c.boxes[i];
        const float mnx = b.center.x - b.halfExtents.x;
        const float mny = b.center.y - b.halfExtents.y;
        const float mnz = b.center.z - b.halfExtents.z;
        const float mxx = b.center.x + b.halfExtents.x;
        const float mxy = b.center.y + b.halfExtents.y;
        const float mxz = b.center.z + b.halfExtents.z;
        const float cx = fminf(fmaxf(p.x, mnx), mxx);
        const float cy = fminf(fmaxf(p.y, mny), mxy);
        const float cz = fminf(fmaxf(p.z, mnz), mxz);
        const float ox = p.x - cx;
        const float oy = p.y - cy;
        const float oz = p.z - cz;
        const float d2 = ox * ox + oy * oy + oz * oz;
        const float rc = r + eps;
        if (d2 > 0.0f && d2 <= rc * rc) {
            const float sc = 1.0f / __builtin_sqrtf(d2);
            visit(ContactGeom{ox * sc, oy * sc, oz * sc, true, sc, ox != 0.0f, oy != 0.0f,
                              oz != 0.0f});
        } else if (d2 == 0.0f) {
            const float dx = fminf(p.x - mnx, mxx - p.x);
            const float dy = fminf(p.y - mny, mxy - p.y);
            const float dz = fminf(p.z - mnz, mxz - p.z);
            float nx = 0.0f, ny = 0.0f, nz = 0.0f;
            if (dx <= dy && dx <= dz)
                nx = p.x - mnx < mxx - p.x ? -1.0f : 1.0f;
            else if (dy <= dz)
                ny = p.y - mny < mxy - p.y ? -1.0f : 1.0f;
            else
                nz = p.z - mnz < mxz - p.z ? -1.0f : 1.0f;
            wall(nx, ny, nz);
        }
    }
    for (int i = 0; i < c.planeCount; ++i) {
        const pbf_plane pl = c.planes[i];
        const float l2 = pl.normal.x * pl.normal.x + pl.normal.y * pl.normal.y +
                         pl.normal.z * pl.normal.z;
        if (l2 == 0.0f) continue;
        const float ns = 1.0f / __builtin_sqrtf(l2);
        const float nx = pl.normal.x * ns;
        const float ny = pl.normal.y * ns;
        const float nz = pl.normal.z * ns;
        const float dist =
            (p.x - pl.point.x) * nx + (p.y - pl.point.y) * ny + (p.z - pl.point.z) * nz;
        if (dist <= r + eps) wall(nx, ny, nz);
    }
}

// resolveVelocitiesKernel body (collision_kernels.cu:236-371).
__device__ __forceinline__ void resolve_velocity(const float4& p, const float4& vin, float4& v,
                                                 const Colliders& c, float r, float restitution,
                                                 float friction) {
#pragma clang fp contract(off)
    const float ts = 1.0f - friction;
    for_each_contact(p, c, r, [&](const ContactGeom& g) {
        contact_velocity(g.nx, g.ny, g.nz, vin, restitution, ts, v);
    });
}

// Neighbour loop: kNbrBatch list entries, then their kNbrBatch records, are
// in flight before the first is used, so a wave's gathers overlap instead of
// waiting one neighbour at a time.  Contributions are still applied in list
// order, so every sum is formed exactly as the one-neighbour-at-a-time loop
// forms it.  kGrouped reads the step's grouped lists (four entries per
// 16-byte load, free slots of the last group repeating its last entry);
// otherwise the reference's slot-major list, where slots past the count
// reload the last valid entry (c >= 1 inside the loop).  Slots past the
// count are not applied.
constexpr int kNbrBatch = 4;  // a group of the step's lists

// kDepth groups (4 * kDepth records) of the grouped lists are in flight per
// step.  The solver step's passes take two, which hides more of the
// gathers' latency at the cost of registers (1M dam break, per launch:
// vorticity 130 -> 110 us, confinement 71 -> 66 us, Delta-p 65.5 -> 63.3 us,
// lambda 65.3 -> 64.5 us); the adjoint's pair kernels gained nothing and
// keep one.
template <bool kGrouped, int kDepth = 1, class Load, class Use>
__device__ __forceinline__ void for_each_neighbour(const uint32_t* __restrict__ nbr, size_t stride,
                                                   size_t i, int c, Load load, Use use) {
    if (kGrouped) {
        // The next groups' entries are loaded while these groups' records
        // are gathered and used: the list read (a stream from HBM) and the
        // record gathers (L2) overlap instead of alternating.
        // The last group is loaded again in place of those past it, so the
        // loads are unconditional.
        constexpr int kStep = kNbrBatch * kDepth;
        const uint4* groups = reinterpret_cast<const uint4*>(nbr) + i;
        const int lastGroup = (c - 1) / 4;
        // A list's stride (the particle capacity) is far below 2^32, so a
        // group's offset is one 32 x 32 -> 64-bit multiply.
        const uint64_t stride32 = static_cast<uint32_t>(stride);
        uint4 g[kDepth];
#pragma unroll
        for (int d = 0; d < kDepth; ++d)
            g[d] = c > 0 ? groups[static_cast<size_t>(min(d, lastGroup)) * stride]
                         : make_uint4(0u, 0u, 0u, 0u);
        for (int q0 = 0; q0 < c; q0 += kStep) {
            uint32_t j[kStep];
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                j[4 * d] = g[d].x;
                j[4 * d + 1] = g[d].y;
                j[4 * d + 2] = g[d].z;
                j[4 * d + 3] = g[d].w;
            }
            decltype(load(0u)) r[kStep];
#pragma unroll
            for (int u = 0; u < kStep; ++u) r[u] = load(j[u]);
#pragma unroll
            for (int d = 0; d < kDepth; ++d)
                g[d] = groups[static_cast<uint64_t>(static_cast<uint32_t>(
                                  min(q0 / 4 + kDepth + d, lastGroup))) *
                              stride32];
#pragma unroll
            for (int u = 0; u < kStep; ++u)
                if (q0 + u < c) use(r[u]);
        }
        return;
    }
    for (int q0 = 0; q0 < c; q0 += kNbrBatch) {
        uint32_t j[kNbrBatch];
        {
#pragma unroll
            for (int u = 0; u < kNbrBatch; ++u)
                j[u] = nbr[static_cast<size_t>(min(q0 + u, c - 1)) * stride + i];
        }
        decltype(load(0u)) r[kNbrBatch];
#pragma unroll
        for (int u = 0; u < kNbrBatch; ++u) r[u] = load(j[u]);
#pragma unroll
        for (int u = 0; u < kNbrBatch; ++u)
            if (q0 + u < c) use(r[u]);
    }
}

// Launch geometry: 256 threads (4 wave64) per block, the reference's kBlockSize.
constexpr int kBlock = 256;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  xcd_block(b, nb) gives every XCD one contiguous run of block
// indices, so a kernel whose neighbouring blocks read overlapping data (the
// cell-sorted working set) keeps that data in one XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned xcd = b & 7u;
    const unsigned k = b >> 3;
    const unsigned per = nb >> 3;
    const unsigned rem = nb & 7u;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

inline unsigned grid_for(size_t n, int block = kBlock) {
    return static_cast<unsigned>((n + static_cast<size_t>(block) - 1) / static_cast<size_t>(block));
}

}  // namespace pbfdev
