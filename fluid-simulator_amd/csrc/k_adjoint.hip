// Adjoint of the PBF substep for gfx950: reverse-mode derivatives of one
// substep (integrate, 4 x [lambda, Delta-p with s_corr, container
// projection], velocity update, XSPH, vorticity, confinement, contact
// velocities) with the neighbour lists and every discrete decision of the
// forward held fixed (which coordinates the container projects, which walls
// a particle touches and which contacts act).
//
// The forward's tape gives, per grid build, the slot -> particle-id map, the
// grouped lists and counts, and the positions of that build's working-set
// order.  Adjoint accumulators of positions and velocities are kept in
// particle-id (API) order; per-build temporaries in that build's slot order.
// A thread owns one slot and loops over its list; what a pair contributes to
// the neighbour goes out by hardware float atomics (the neighbour's own
// thread could pull it instead, at the price of a second loop over the
// same pairs).  Sums therefore run in no fixed order: the adjoint is
// deterministic only to rounding.
//
// Reference of the differentiated forward: src/pbf/pbf_solver.cu:186-328,
// src/pbf/pbf_solver_kernels.cu:26-385, src/pbf/collision_kernels.cu:21-82,
// :236-282, include/pbf/sph_helpers.cuh:13-95.
#include "pbf_device.hpp"
#include "pbf_internal.hpp"
#include "pbf_launch.hpp"

using namespace pbfdev;

namespace {

__device__ __forceinline__ float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) {
    return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ float3 operator-(float3 a, float3 b) {
    return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float3 operator*(float s, float3 a) {
    return make_float3(s * a.x, s * a.y, s * a.z);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ void atomic_add3(pbf_float4* base, uint32_t i, float3 v) {
    float* p = reinterpret_cast<float*>(base + i);
    unsafeAtomicAdd(p + 0, v.x);
    unsafeAtomicAdd(p + 1, v.y);
    unsafeAtomicAdd(p + 2, v.z);
}

// Sum over the wave, then one double atomic per wave.
__device__ __forceinline__ void wave_add(double* acc, float v) {
    double s = static_cast<double>(v);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(acc, s);
}

// G(d) = spiky gradient (correctly rounded form of the forward).
__device__ __forceinline__ float3 spiky(float3 d, const Sph& s) {
    const float3 g = spiky_grad(d.x, d.y, d.z, dot3(d, d), s);
    return g;
}

// J_G(d) a: the Jacobian of the spiky gradient (symmetric) applied to a.
// With u = d / r and w = 1 - r / h, G = A w^2 u and
// J = A [ w^2 (I - u u^T) / r - (2 / h) w u u^T ].
__device__ __forceinline__ float3 spiky_jac(float3 d, float3 a, const Sph& s) {
    const float r = __builtin_sqrtf(dot3(d, d));
    if (r <= 0.0f || r >= s.h) return make_float3(0.0f, 0.0f, 0.0f);
    const float inv_r = 1.0f / r;
    const float3 u = inv_r * d;
    const float w = (s.h - r) * s.inv_h;
    const float ua = dot3(u, a);
    return s.spiky_norm * ((w * w * inv_r) * (a - ua * u) - (2.0f * s.inv_h * w * ua) * u);
}

// Gradient of the poly6 kernel with respect to d.
__device__ __forceinline__ float3 poly6_grad(float3 d, const Sph& s) {
    const float d2 = dot3(d, d);
    if (d2 >= s.h2) return make_float3(0.0f, 0.0f, 0.0f);
    const float q = (s.h2 - d2) * s.inv_h2;
    return (s.poly_norm * 3.0f * q * q * (-2.0f * s.inv_h2)) * d;
}

// Gradient of the artificial pressure -k ((h^2 - d^2) / (h^2 - dq^2))^(3n).
__device__ __forceinline__ float3 scorr_grad(float3 d, float h2, const Scorr& c) {
    const float d2 = dot3(d, d);
    if (!c.enabled || d2 >= h2) return make_float3(0.0f, 0.0f, 0.0f);
    const float q = (h2 - d2) * c.inv_den;
    const int k3n = 3 * c.n;
    return (-c.k * static_cast<float>(k3n) * ipow(q, k3n - 1) * (-2.0f * c.inv_den)) * d;
}

// A neighbour's index and two of its records, loaded together.
struct Nbr2 {
    uint32_t j;
    float4 a, b;
};

struct Walls {
    float minX, maxX, minY, minZ, maxZ;
};

__device__ __forceinline__ Walls walls_of(const pbf_container& k, float r) {
    return Walls{k.min.x + r, k.max.x - r, k.min.y + r, k.min.z + r, k.max.z - r};
}

// ---------------------------------------------------------------------------
// Post-solve phase (one build: the post-solve rebuild)
// ---------------------------------------------------------------------------
struct PostArgs {
    pbfint::AdjBuild b;
    const pbf_float4* xs;   // substep start positions
    const pbf_float4* pf;   // final positions
    const pbf_float4* v3;   // velocities after XSPH
    const pbf_float4* om;   // vorticity, |omega| in .w
    const pbf_float4* vin;  // velocities after gravity (contact inputs)
    size_t n;
    float h, dt, visc, vort, radius, restitution, friction;
    pbf_container box;
};

// Confinement and contact velocities: recomputes eta, N and v4, replays the
// wall sequence of resolveVelocities, and reverses both.
__global__ void __launch_bounds__(kBlock)
    k_adj_post1(PostArgs a, const pbf_float4* __restrict__ vb, pbf_float4* __restrict__ v1b,
                pbf_float4* __restrict__ v3b, pbf_float4* __restrict__ etab,
                pbf_float4* __restrict__ wb, float* __restrict__ wlb, double* __restrict__ acc) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float epart = 0.0f;
    if (i < a.n) {
        const Sph s = sph_consts(a.h);
        const uint32_t gid = a.b.ids[i];
        const float3 pi = f3(ld4(a.pf, i));
        const float4 wi = ld4(a.om, i);
        const float3 w = f3(wi);
        float3 eta = make_float3(0.0f, 0.0f, 0.0f);
        for_each_neighbour<true>(
            a.b.nbr, a.b.stride, i, a.b.cnt[i],
            [&](uint32_t j) { return Nbr2{j, ld4(a.pf, j), ld4(a.om, j)}; },
            [&](const Nbr2& r) { eta = eta + (r.b.w - wi.w) * spiky(pi - f3(r.a), s); });
        const float el = __builtin_sqrtf(dot3(eta, eta));
        const bool ok = el > 1.0e-6f;
        const float3 N = ok ? (1.0f / el) * eta : make_float3(0.0f, 0.0f, 0.0f);
        const float3 v3i = f3(ld4(a.v3, i));
        const float3 v4 = v3i + (a.dt * a.vort) * cross3(N, w);
        // Replay the walls of resolveVelocities on v4.
        const Walls wl = walls_of(a.box, a.radius);
        constexpr float eps = 1e-4f;
        const float3 vi1 = f3(ld4(a.vin, i));
        const float3 nrm[5] = {make_float3(1, 0, 0), make_float3(-1, 0, 0), make_float3(0, 1, 0),
                               make_float3(0, 0, 1), make_float3(0, 0, -1)};
        bool touch[5];
        touch[0] = pi.x <= wl.minX + eps;
        touch[1] = !touch[0] && pi.x >= wl.maxX - eps;
        touch[2] = pi.y <= wl.minY + eps;
        touch[3] = pi.z <= wl.minZ + eps;
        touch[4] = !touch[3] && pi.z >= wl.maxZ - eps;
        const float ts = 1.0f - a.friction;
        bool act[5], inNeg[5];
        float3 v = v4;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const float vn = dot3(v, nrm[k]);
            const float vinn = dot3(vi1, nrm[k]);
            const float target = vinn < 0.0f ? -a.restitution * vinn : 0.0f;
            inNeg[k] = vinn < 0.0f;
            act[k] = touch[k] && !(vn > target || (vn == target && vinn >= 0.0f));
            if (act[k]) v = target * nrm[k] + ts * (v - vn * nrm[k]);
        }
        float3 vc = f3(ld4(vb, gid));
        float3 b1 = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int k = 4; k >= 0; --k) {
            if (!act[k]) continue;
            const float vnb = dot3(vc, nrm[k]);
            if (inNeg[k]) b1 = b1 + (-a.restitution * vnb) * nrm[k];
            vc = ts * (vc - vnb * nrm[k]);
        }
        st4(v1b, gid, make_float4(b1.x, b1.y, b1.z, 0.0f));
        // confinement: v4 = v3 + dt e (N x w)
        st4(v3b, i, make_float4(vc.x, vc.y, vc.z, 0.0f));
        epart = a.dt * dot3(cross3(N, w), vc);
        const float3 cb = (a.dt * a.vort) * vc;
        const float3 Nb = cross3(w, cb);
        const float3 wbar = cross3(cb, N);
        const float3 eb = ok ? (1.0f / el) * (Nb - dot3(N, Nb) * N) : make_float3(0.0f, 0.0f, 0.0f);
        st4(etab, i, make_float4(eb.x, eb.y, eb.z, 0.0f));
        st4(wb, i, make_float4(wbar.x, wbar.y, wbar.z, 0.0f));
        wlb[i] = 0.0f;
    }
    wave_add(acc + 1, epart);
}

// eta_i = sum_j (|w_j| - |w_i|) G(d_ij): adjoints of |w| and of positions.
__global__ void __launch_bounds__(kBlock)
    k_adj_post2(PostArgs a, const pbf_float4* __restrict__ etab, float* __restrict__ wlb,
                pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float3 pi = f3(ld4(a.pf, i));
    const float wli = a.om[i].w;
    const float3 eb = f3(ld4(etab, i));
    float wacc = 0.0f;
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    const uint32_t* ids = a.b.ids;
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) { return Nbr2{j, ld4(a.pf, j), ld4(a.om, j)}; },
        [&](const Nbr2& r) {
            const uint32_t j = r.j;
            const float3 d = pi - f3(r.a);
            const float3 G = spiky(d, s);
            const float gdot = dot3(eb, G);
            unsafeAtomicAdd(wlb + j, gdot);
            wacc -= gdot;
            const float3 db = (r.b.w - wli) * spiky_jac(d, eb, s);
            pacc = pacc + db;
            atomic_add3(pb, ids[j], make_float3(-db.x, -db.y, -db.z));
        });
    unsafeAtomicAdd(wlb + i, wacc);
    atomic_add3(pb, ids[i], pacc);
}

// |w|-bar into w-bar; clears the XSPH accumulator.
__global__ void __launch_bounds__(kBlock)
    k_adj_post3(PostArgs a, const float* __restrict__ wlb, pbf_float4* __restrict__ wb,
                pbf_float4* __restrict__ v2b) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float4 wi = ld4(a.om, i);
    float3 w = f3(ld4(wb, i));
    if (wi.w > 0.0f) w = w + (wlb[i] / wi.w) * f3(wi);
    st4(wb, i, make_float4(w.x, w.y, w.z, 0.0f));
    st4(v2b, i, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
}

// w_i = sum_j G(d_ij) x (v3_j - v3_i).
__global__ void __launch_bounds__(kBlock)
    k_adj_post4(PostArgs a, const pbf_float4* __restrict__ wb, pbf_float4* __restrict__ v3b,
                pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float3 pi = f3(ld4(a.pf, i));
    const float3 vi = f3(ld4(a.v3, i));
    const float3 wbar = f3(ld4(wb, i));
    float3 vacc = make_float3(0.0f, 0.0f, 0.0f);
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    const uint32_t* ids = a.b.ids;
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) { return Nbr2{j, ld4(a.pf, j), ld4(a.v3, j)}; },
        [&](const Nbr2& r) {
            const uint32_t j = r.j;
            const float3 d = pi - f3(r.a);
            const float3 G = spiky(d, s);
            const float3 u = f3(r.b) - vi;
            const float3 ub = cross3(wbar, G);
            atomic_add3(v3b, j, ub);
            vacc = vacc - ub;
            const float3 db = spiky_jac(d, cross3(u, wbar), s);
            pacc = pacc + db;
            atomic_add3(pb, ids[j], make_float3(-db.x, -db.y, -db.z));
        });
    atomic_add3(v3b, static_cast<uint32_t>(i), vacc);
    atomic_add3(pb, ids[i], pacc);
}

// v3_i = v2_i + c sum_j (v2_j - v2_i) W(d_ij), v2 = (p - x) / dt.
__global__ void __launch_bounds__(kBlock)
    k_adj_post5(PostArgs a, const pbf_float4* __restrict__ v3b, pbf_float4* __restrict__ v2b,
                pbf_float4* __restrict__ pb, double* __restrict__ acc) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float cpart = 0.0f;
    if (i < a.n) {
        const Sph s = sph_consts(a.h);
        const float inv_dt = 1.0f / a.dt;
        const float3 pi = f3(ld4(a.pf, i));
        const float3 v2i = inv_dt * (pi - f3(ld4(a.xs, i)));
        const float3 vb3 = f3(ld4(v3b, i));
        float3 vacc = vb3;
        float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
        const uint32_t* ids = a.b.ids;
        for_each_neighbour<true>(
            a.b.nbr, a.b.stride, i, a.b.cnt[i],
            [&](uint32_t j) { return Nbr2{j, ld4(a.pf, j), ld4(a.xs, j)}; },
            [&](const Nbr2& r) {
                const uint32_t j = r.j;
                const float3 pj = f3(r.a);
                const float3 d = pi - pj;
                const float W = poly6_d2(dot3(d, d), s);
                const float3 dv = inv_dt * (pj - f3(r.b)) - v2i;
                const float3 t = (a.visc * W) * vb3;
                atomic_add3(v2b, j, t);
                vacc = vacc - t;
                cpart += W * dot3(vb3, dv);
                const float3 db = (a.visc * dot3(dv, vb3)) * poly6_grad(d, s);
                pacc = pacc + db;
                atomic_add3(pb, ids[j], make_float3(-db.x, -db.y, -db.z));
            });
        atomic_add3(v2b, static_cast<uint32_t>(i), vacc);
        atomic_add3(pb, ids[i], pacc);
    }
    wave_add(acc + 0, cpart);
}

// v2 = (p - x) / dt: into the position adjoint and the start positions'.
__global__ void __launch_bounds__(kBlock)
    k_adj_post6(PostArgs a, const pbf_float4* __restrict__ v2b, pbf_float4* __restrict__ pb,
                pbf_float4* __restrict__ x0b) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t gid = a.b.ids[i];
    const float inv_dt = 1.0f / a.dt;
    const float3 g = inv_dt * f3(ld4(v2b, i));
    float4 p = ld4(pb, gid);
    p.x += g.x;
    p.y += g.y;
    p.z += g.z;
    st4(pb, gid, p);
    st4(x0b, gid, make_float4(-g.x, -g.y, -g.z, 0.0f));
}

// ---------------------------------------------------------------------------
// Constraint iteration (one build)
// ---------------------------------------------------------------------------
struct IterArgs {
    pbfint::AdjBuild b;
    const pbf_float4* pk;    // positions entering the iteration, lambda in .w
    const pbf_float4* pout;  // positions leaving it (after the projection)
    size_t n;
    float h, mass, rho0, eps, sk, sdq, radius;
    int sn;
    pbf_container box;
};

// The projection's mask, then the lambda terms (density and gradient sums
// in FP64, as the forward) for the next two passes.
__global__ void __launch_bounds__(kBlock)
    k_adj_it1(IterArgs a, pbf_float4* __restrict__ pb, pbf_float4* __restrict__ prb,
              pbf_float4* __restrict__ gsum, float* __restrict__ cterm,
              float* __restrict__ lamb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t gid = a.b.ids[i];
    const Walls wl = walls_of(a.box, a.radius);
    const float4 po = ld4(a.pout, i);
    // A coordinate the container moved sits exactly on its wall.
    const bool kx = !(po.x == wl.minX || po.x == wl.maxX);
    const bool ky = !(po.y == wl.minY);
    const bool kz = !(po.z == wl.minZ || po.z == wl.maxZ);
    const float4 p = ld4(pb, gid);
    const float4 r = make_float4(kx ? p.x : 0.0f, ky ? p.y : 0.0f, kz ? p.z : 0.0f, 0.0f);
    st4(pb, gid, r);
    st4(prb, i, r);
    const Sph s = sph_consts(a.h);
    const float gs = a.mass / a.rho0;
    const float3 pi = f3(ld4(a.pk, i));
    double rho = static_cast<double>(a.mass * s.poly_norm);
    double sg = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i], [&](uint32_t j) { return ld4(a.pk, j); },
        [&](const float4& pj) {
            const float3 d = pi - f3(pj);
            rho += static_cast<double>(a.mass * poly6_d2(dot3(d, d), s));
            const float3 g = (-gs) * spiky(d, s);
            sg += static_cast<double>(dot3(g, g));
            gx += g.x;
            gy += g.y;
            gz += g.z;
        });
    sg += gx * gx + gy * gy + gz * gz;
    st4(gsum, i, make_float4(static_cast<float>(gx), static_cast<float>(gy),
                             static_cast<float>(gz), static_cast<float>(sg)));
    cterm[i] = static_cast<float>(rho / static_cast<double>(a.rho0) - 1.0);
    lamb[i] = 0.0f;
}

// Delta-p_i = (m / rho0) sum_j (l_i + l_j + s_ij) G(d_ij).
__global__ void __launch_bounds__(kBlock)
    k_adj_it2(IterArgs a, const pbf_float4* __restrict__ prb, float* __restrict__ lamb,
              pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const Scorr sc = scorr_consts(a.h, a.sk, a.sn, a.sdq);
    const float gs = a.mass / a.rho0;
    const float4 pk = ld4(a.pk, i);
    const float3 pi = f3(pk);
    const float3 r = f3(ld4(prb, i));
    float lacc = 0.0f;
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    const uint32_t* ids = a.b.ids;
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i], [&](uint32_t j) { return Nbr2{j, ld4(a.pk, j), {}}; },
        [&](const Nbr2& rr) {
            const uint32_t j = rr.j;
            const float4 pj = rr.a;
            const float3 d = pi - f3(pj);
            const float d2 = dot3(d, d);
            const float3 G = spiky(d, s);
            const float coef = pk.w + pj.w + scorr_value(d2, s.h2, sc);
            const float q = gs * dot3(r, G);
            lacc += q;
            unsafeAtomicAdd(lamb + j, q);
            const float3 db = (gs * coef) * spiky_jac(d, r, s) + q * scorr_grad(d, s.h2, sc);
            pacc = pacc + db;
            atomic_add3(pb, ids[j], make_float3(-db.x, -db.y, -db.z));
        });
    unsafeAtomicAdd(lamb + i, lacc);
    atomic_add3(pb, ids[i], pacc);
}

// lambda_i = -C_i / (S_i + eps).
__global__ void __launch_bounds__(kBlock)
    k_adj_it3(IterArgs a, const pbf_float4* __restrict__ gsum, const float* __restrict__ cterm,
              const float* __restrict__ lamb, pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float gs = a.mass / a.rho0;
    const float4 gsi = ld4(gsum, i);
    const float den = gsi.w + a.eps;
    const float lb = lamb[i];
    const float Cb = -lb / den;
    const float Sb = lb * cterm[i] / (den * den);
    const float3 pi = f3(ld4(a.pk, i));
    const float3 gsv = f3(gsi);
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    const uint32_t* ids = a.b.ids;
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i], [&](uint32_t j) { return Nbr2{j, ld4(a.pk, j), {}}; },
        [&](const Nbr2& rr) {
            const uint32_t j = rr.j;
            const float3 d = pi - f3(rr.a);
            const float3 g = (-gs) * spiky(d, s);
            const float3 gb = (2.0f * Sb) * (g + gsv);
            const float3 db = (Cb * gs) * poly6_grad(d, s) + (-gs) * spiky_jac(d, gb, s);
            pacc = pacc + db;
            atomic_add3(pb, ids[j], make_float3(-db.x, -db.y, -db.z));
        });
    atomic_add3(pb, ids[i], pacc);
}

// Integrate: p0 = x + v1 dt, v1 = v + g dt.  In id order.
__global__ void __launch_bounds__(kBlock)
    k_adj_predict(const pbf_float4* __restrict__ pb, pbf_float4* __restrict__ x0b,
                  pbf_float4* __restrict__ v1b, size_t n, float dt) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = ld4(pb, i);
    float4 x = ld4(x0b, i);
    float4 v = ld4(v1b, i);
    x.x += p.x;
    x.y += p.y;
    x.z += p.z;
    v.x += p.x * dt;
    v.y += p.y * dt;
    v.z += p.z * dt;
    st4(x0b, i, x);
    st4(v1b, i, v);
}

}  // namespace

namespace pbfint {

pbf_status adjoint_post(const AdjPost& t, const AdjBuffers& w, size_t n, const AdjParams& prm,
                        void* stream) {
    if (n == 0) return PBF_OK;
    PostArgs a;
    a.b = t.build;
    a.xs = t.xs;
    a.pf = t.pf;
    a.v3 = t.v3;
    a.om = t.omega;
    a.vin = t.vin;
    a.n = n;
    a.h = prm.h;
    a.dt = prm.dt;
    a.visc = prm.visc;
    a.vort = prm.vort;
    a.radius = prm.radius;
    a.restitution = prm.restitution;
    a.friction = prm.friction;
    a.box = prm.box;
    const unsigned g = grid_for(n);
    pbf_status st;
    hipLaunchKernelGGL(k_adj_post1, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.vb, w.v1b, w.v3b, w.etab, w.wb, w.wlb, w.acc);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    hipLaunchKernelGGL(k_adj_post2, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.etab, w.wlb, w.pb);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    hipLaunchKernelGGL(k_adj_post3, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.wlb, w.wb, w.v2b);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    hipLaunchKernelGGL(k_adj_post4, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.wb, w.v3b, w.pb);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    hipLaunchKernelGGL(k_adj_post5, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.v3b, w.v2b, w.pb, w.acc);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    PBF_LAUNCH(k_adj_post6, g, kBlock, stream, a, w.v2b, w.pb, w.x0b);
}

pbf_status adjoint_iteration(const AdjIter& t, const AdjBuffers& w, size_t n, const AdjParams& prm,
                             void* stream) {
    if (n == 0) return PBF_OK;
    IterArgs a;
    a.b = t.build;
    a.pk = t.pk;
    a.pout = t.pout;
    a.n = n;
    a.h = prm.h;
    a.mass = prm.mass;
    a.rho0 = prm.rho0;
    a.eps = prm.lambdaEps;
    a.sk = prm.scorrK;
    a.sn = prm.scorrN;
    a.sdq = prm.scorrDq;
    a.radius = prm.radius;
    a.box = prm.box;
    const unsigned g = grid_for(n);
    pbf_status st;
    hipLaunchKernelGGL(k_adj_it1, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.pb, w.prb, w.gsum, w.cterm, w.lamb);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    hipLaunchKernelGGL(k_adj_it2, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.prb, w.lamb, w.pb);
    if ((st = pbf_hip_check(hipGetLastError())) != PBF_OK) return st;
    PBF_LAUNCH(k_adj_it3, g, kBlock, stream, a, w.gsum, w.cterm, w.lamb, w.pb);
}

pbf_status adjoint_predict(const AdjBuffers& w, size_t n, float dt, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_adj_predict, grid_for(n), kBlock, stream, w.pb, w.x0b, w.v1b, n, dt);
}

}  // namespace pbfint
