// Adjoint of the PBF substep for gfx950: reverse-mode derivatives of one
// substep (integrate, the constraint iterations [lambda, Delta-p with
// s_corr, the collision solve against the container, spheres, rounded boxes
// and planes], velocity update, XSPH, vorticity, confinement, contact
// velocities against every collider) with the neighbour lists and every
// discrete decision of the forward held fixed (which projections act and
// their branches, which contacts touch and act).
//
// The forward's tape gives, per grid build, the slot -> particle-id map, the
// grouped lists and counts, and the positions of that build's working-set
// order.  Adjoint accumulators of positions and velocities are kept in
// particle-id (API) order; per-build temporaries in that build's slot order.
// A thread owns one slot and loops over its list once per pass, pulling
// both what its own pair terms owe it and what its neighbours' terms owe it:
// the lists are symmetric (j lists i exactly when i lists j), G is odd in
// d and its Jacobian even, so the neighbour's term is a function of the
// neighbour's per-slot records and the same d.  No pair term goes through an
// atomic; only the two parameter sums do (per wave, in FP64).
//
// Reference of the differentiated forward: src/pbf/pbf_solver.cu:186-328,
// src/pbf/pbf_solver_kernels.cu:26-385, src/pbf/collision_kernels.cu:7-371,
// src/pbf/collision_system.cu:133-218, include/pbf/sph_helpers.cuh:13-95.
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
// Three per-slot records of a neighbour, loaded together with the group's
// other entries' (for_each_neighbour keeps a group's loads in flight).
struct Nbr3 {
    uint32_t j;
    float4 a, b, c;
};

// A 3x3 matrix by rows.
struct M3 {
    float3 r0, r1, r2;
};

__device__ __forceinline__ M3 m3_identity() {
    return M3{make_float3(1.0f, 0.0f, 0.0f), make_float3(0.0f, 1.0f, 0.0f),
              make_float3(0.0f, 0.0f, 1.0f)};
}

// A B.
__device__ __forceinline__ M3 m3_mul(const M3& A, const M3& B) {
    const auto row = [&](float3 a) {
        return make_float3(a.x * B.r0.x + a.y * B.r1.x + a.z * B.r2.x,
                           a.x * B.r0.y + a.y * B.r1.y + a.z * B.r2.y,
                           a.x * B.r0.z + a.y * B.r1.z + a.z * B.r2.z);
    };
    return M3{row(A.r0), row(A.r1), row(A.r2)};
}

// A^T b.
__device__ __forceinline__ float3 m3_tmul(const M3& A, float3 b) {
    return b.x * A.r0 + b.y * A.r1 + b.z * A.r2;
}

// scale (I - u u^T).
__device__ __forceinline__ M3 radial_jac(float3 u, float scale) {
    return M3{scale * (make_float3(1.0f, 0.0f, 0.0f) - u.x * u),
              scale * (make_float3(0.0f, 1.0f, 0.0f) - u.y * u),
              scale * (make_float3(0.0f, 0.0f, 1.0f) - u.z * u)};
}

// Zeroes the rows of J whose coordinate differs between a and b (a
// coordinate a projection set to a constant).
__device__ __forceinline__ void zero_changed_rows(M3& J, const float4& a, const float4& b) {
    if (a.x != b.x) J.r0 = make_float3(0.0f, 0.0f, 0.0f);
    if (a.y != b.y) J.r1 = make_float3(0.0f, 0.0f, 0.0f);
    if (a.z != b.z) J.r2 = make_float3(0.0f, 0.0f, 0.0f);
}

// The collision fixed point of CollisionSystem::solve (collision_fixed_point)
// replayed from q with the Jacobian of every projection that acts
// accumulated: J = J_m ... J_1 for the m projections that move q, each with
// its decisions (which projection acts, the sphere's centre case, the box's
// clamped axes or interior face) held fixed.
//   container: the clamped coordinates' rows vanish;
//   sphere, q' = c + (md / |o|) o: (md / |o|) (I - u u^T); at o = 0 the x
//     row vanishes;
//   box outside its core, q' = cp + (r / |o|) o, cp = clamp(q):
//     diag(inside) + (r / |o|) (I - u u^T) diag(outside); inside the core the
//     face coordinate's row vanishes;
//   plane, q' = q + n (r - n.(q - p0)): I - n n^T.
__device__ __forceinline__ M3 collide_jacobian(float4 q, const Colliders& c, float r) {
    M3 J = m3_identity();
    for (int pass = 0; pass < kMaxCollisionPasses; ++pass) {
        bool marked = false;
        float4 before = q;
        if (project_container(q, *c.container, r)) {
            zero_changed_rows(J, before, q);
            marked = true;
        }
        for (int i = 0; i < c.sphereCount; ++i) {
            before = q;
            const pbf_sphere sp = c.spheres[i];
            if (!project_sphere(q, sp, r)) continue;
            marked = true;
            const float3 o = make_float3(before.x - sp.center.x, before.y - sp.center.y,
                                         before.z - sp.center.z);
            const float d2 = dot3(o, o);
            if (d2 == 0.0f) {
                J.r0 = make_float3(0.0f, 0.0f, 0.0f);
            } else {
                const float inv = 1.0f / __builtin_sqrtf(d2);
                J = m3_mul(radial_jac(inv * o, (sp.radius + r) * inv), J);
            }
        }
        for (int i = 0; i < c.boxCount; ++i) {
            before = q;
            const pbf_box b = c.boxes[i];
            if (!project_box(q, b, r)) continue;
            marked = true;
            const float3 mn = make_float3(b.center.x - b.halfExtents.x,
                                          b.center.y - b.halfExtents.y,
                                          b.center.z - b.halfExtents.z);
            const float3 mx = make_float3(b.center.x + b.halfExtents.x,
                                          b.center.y + b.halfExtents.y,
                                          b.center.z + b.halfExtents.z);
            const float3 cp = make_float3(fminf(fmaxf(before.x, mn.x), mx.x),
                                          fminf(fmaxf(before.y, mn.y), mx.y),
                                          fminf(fmaxf(before.z, mn.z), mx.z));
            const float3 o = f3(before) - cp;
            const float d2 = dot3(o, o);
            if (d2 == 0.0f) {
                zero_changed_rows(J, before, q);
            } else {
                const float inv = 1.0f / __builtin_sqrtf(d2);
                M3 Jk = radial_jac(inv * o, r * inv);
                // columns of the clamped (outside) axes only, plus the
                // identity on the axes the clamp passes through
                const bool inX = o.x == 0.0f, inY = o.y == 0.0f, inZ = o.z == 0.0f;
                const auto mask = [&](float3 row, float3 unit) {
                    return make_float3(inX ? unit.x : row.x, inY ? unit.y : row.y,
                                       inZ ? unit.z : row.z);
                };
                Jk.r0 = mask(Jk.r0, make_float3(1.0f, 0.0f, 0.0f));
                Jk.r1 = mask(Jk.r1, make_float3(0.0f, 1.0f, 0.0f));
                Jk.r2 = mask(Jk.r2, make_float3(0.0f, 0.0f, 1.0f));
                J = m3_mul(Jk, J);
            }
        }
        for (int i = 0; i < c.planeCount; ++i) {
            const pbf_plane pl = c.planes[i];
            if (!project_plane(q, pl, r)) continue;
            marked = true;
            const float l2 = dot3(make_float3(pl.normal.x, pl.normal.y, pl.normal.z),
                                  make_float3(pl.normal.x, pl.normal.y, pl.normal.z));
            const float ns = 1.0f / __builtin_sqrtf(l2);
            const float3 nu = make_float3(pl.normal.x * ns, pl.normal.y * ns, pl.normal.z * ns);
            J = m3_mul(radial_jac(nu, 1.0f), J);
        }
        if (!marked) break;
    }
    return J;
}

// One acting contact of resolveVelocitiesKernel as the reverse pass needs
// it: its geometry, target normal speed, whether the incoming velocity
// points into the surface, and the velocity entering it.
struct ActingContact {
    ContactGeom g;
    float target;
    bool neg;
    float3 v;
};

// Replays the contacts on v (resolve_velocity's arithmetic, contact by
// contact) and counts those that act.  With stop >= 0 it returns as soon as
// acting contact number `stop` is reached, without applying it, and
// describes it in *out.  Returns the number of acting contacts applied.
__device__ __forceinline__ int replay_contacts(const float4& p, const float4& vin, float4& v,
                                               const Colliders& c, float r, float restitution,
                                               float friction, int stop, ActingContact* out) {
#pragma clang fp contract(off)
    const float ts = 1.0f - friction;
    int count = 0;
    bool done = false;
    for_each_contact(p, c, r, [&](const ContactGeom& g) {
        if (done) return;
        const float vn = v.x * g.nx + v.y * g.ny + v.z * g.nz;
        const float vinn = vin.x * g.nx + vin.y * g.ny + vin.z * g.nz;
        const float target = vinn < 0.0f ? -restitution * vinn : 0.0f;
        if (vn > target || (vn == target && vinn >= 0.0f)) return;
        if (count == stop) {
            *out = ActingContact{g, target, vinn < 0.0f, f3(v)};
            done = true;
            return;
        }
        contact_velocity(g.nx, g.ny, g.nz, vin, restitution, ts, v);
        ++count;
    });
    return count;
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
    Colliders col;
};

// Confinement and contact velocities: recomputes eta, N and v4, replays the
// contact sequence of resolveVelocities (walls, spheres, boxes, planes),
// and reverses both.  A contact maps v to t n + ts (v - (v.n) n) with
// t = -e (v_in.n) when v_in.n < 0; the reverse of acting contact m needs
// the velocity entering it, which a replay of the first m contacts gives
// (contacts are few per particle).  Sphere and box normals follow the
// position, n = o / |o|, so those contacts also add to the position's
// adjoint: (I - n n^T) n-bar / |o| on the axes o follows.
__global__ void __launch_bounds__(kBlock)
    k_adj_post1(PostArgs a, const pbf_float4* __restrict__ vb, pbf_float4* __restrict__ v1b,
                pbf_float4* __restrict__ v3b, pbf_float4* __restrict__ etab,
                pbf_float4* __restrict__ wb, pbf_float4* __restrict__ pb,
                double* __restrict__ acc) {
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
        const float4 pf4 = ld4(a.pf, i);
        const float4 vin4 = ld4(a.vin, i);
        const float3 vi1 = f3(vin4);
        const float4 v4w = make_float4(v4.x, v4.y, v4.z, 0.0f);
        float4 vtmp = v4w;
        const int acting = replay_contacts(pf4, vin4, vtmp, a.col, a.radius, a.restitution,
                                           a.friction, -1, nullptr);
        const float ts = 1.0f - a.friction;
        float3 vc = f3(ld4(vb, gid));
        float3 b1 = make_float3(0.0f, 0.0f, 0.0f);
        float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
        for (int m = acting; m-- > 0;) {
            ActingContact ac;
            vtmp = v4w;
            replay_contacts(pf4, vin4, vtmp, a.col, a.radius, a.restitution, a.friction, m, &ac);
            const float3 nrm = make_float3(ac.g.nx, ac.g.ny, ac.g.nz);
            const float vnb = dot3(vc, nrm);  // adjoint of the target speed
            if (ac.neg) b1 = b1 + (-a.restitution * vnb) * nrm;
            if (ac.g.radial) {
                const float3 nb = (ac.neg ? -a.restitution * vnb : 0.0f) * vi1 + ac.target * vc -
                                  (ts * vnb) * ac.v - (ts * dot3(ac.v, nrm)) * vc;
                const float3 g = ac.g.inv_len * (nb - dot3(nrm, nb) * nrm);
                pacc = pacc + make_float3(ac.g.outX ? g.x : 0.0f, ac.g.outY ? g.y : 0.0f,
                                          ac.g.outZ ? g.z : 0.0f);
            }
            vc = ts * (vc - vnb * nrm);
        }
        float4 pbi = ld4(pb, gid);
        pbi.x += pacc.x;
        pbi.y += pacc.y;
        pbi.z += pacc.z;
        st4(pb, gid, pbi);
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
    }
    wave_add(acc + 1, epart);
}

// eta_i = sum_j (|w_j| - |w_i|) G(d_ij).  Pulled: |w|-bar_i = -sum_j
// (eta-bar_i + eta-bar_j) . G_ij, and d-bar_ij - d-bar_ji = a_ij J (eta-bar_i +
// eta-bar_j); then |w|-bar into w-bar.
__global__ void __launch_bounds__(kBlock)
    k_adj_post2(PostArgs a, const pbf_float4* __restrict__ etab, pbf_float4* __restrict__ wb,
                pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float3 pi = f3(ld4(a.pf, i));
    const float4 wi = ld4(a.om, i);
    const float3 ebi = f3(ld4(etab, i));
    float wl = 0.0f;
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) { return Nbr3{j, ld4(a.pf, j), ld4(a.om, j), ld4(etab, j)}; },
        [&](const Nbr3& r) {
            const float3 d = pi - f3(r.a);
            const float3 eb = ebi + f3(r.c);
            wl -= dot3(eb, spiky(d, s));
            pacc = pacc + (r.b.w - wi.w) * spiky_jac(d, eb, s);
        });
    float3 w = f3(ld4(wb, i));
    if (wi.w > 0.0f) w = w + (wl / wi.w) * f3(wi);
    st4(wb, i, make_float4(w.x, w.y, w.z, 0.0f));
    const uint32_t gid = a.b.ids[i];
    float4 p = ld4(pb, gid);
    p.x += pacc.x;
    p.y += pacc.y;
    p.z += pacc.z;
    st4(pb, gid, p);
}

// w_i = sum_j G(d_ij) x (v3_j - v3_i).  Pulled: v3-bar_i -= sum_j (w-bar_i +
// w-bar_j) x G_ij, and d-bar_ij - d-bar_ji = J (u_ij x (w-bar_i + w-bar_j)).
__global__ void __launch_bounds__(kBlock)
    k_adj_post4(PostArgs a, const pbf_float4* __restrict__ wb, pbf_float4* __restrict__ v3b,
                pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float3 pi = f3(ld4(a.pf, i));
    const float3 vi = f3(ld4(a.v3, i));
    const float3 wbi = f3(ld4(wb, i));
    float3 vacc = f3(ld4(v3b, i));
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) { return Nbr3{j, ld4(a.pf, j), ld4(a.v3, j), ld4(wb, j)}; },
        [&](const Nbr3& r) {
            const float3 d = pi - f3(r.a);
            const float3 wbs = wbi + f3(r.c);
            vacc = vacc - cross3(wbs, spiky(d, s));
            pacc = pacc + spiky_jac(d, cross3(f3(r.b) - vi, wbs), s);
        });
    st4(v3b, i, make_float4(vacc.x, vacc.y, vacc.z, 0.0f));
    const uint32_t gid = a.b.ids[i];
    float4 p = ld4(pb, gid);
    p.x += pacc.x;
    p.y += pacc.y;
    p.z += pacc.z;
    st4(pb, gid, p);
}

// v3_i = v2_i + c sum_j (v2_j - v2_i) W_ij with v2 = (p - x) / dt.  Pulled:
// v2-bar_i = v3-bar_i + c sum_j W_ij (v3-bar_j - v3-bar_i) and d-bar_ij -
// d-bar_ji = c ((v2_j - v2_i) . (v3-bar_i - v3-bar_j)) grad W_ij; then v2's
// adjoint into the positions' (+1/dt) and the start positions' (-1/dt).
__global__ void __launch_bounds__(kBlock)
    k_adj_post5(PostArgs a, const pbf_float4* __restrict__ v3b, pbf_float4* __restrict__ pb,
                pbf_float4* __restrict__ x0b, double* __restrict__ acc) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float cpart = 0.0f;
    if (i < a.n) {
        const Sph s = sph_consts(a.h);
        const float inv_dt = 1.0f / a.dt;
        const float3 pi = f3(ld4(a.pf, i));
        const float3 v2i = inv_dt * (pi - f3(ld4(a.xs, i)));
        const float3 vbi = f3(ld4(v3b, i));
        float3 v2b = vbi;
        float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
        for_each_neighbour<true>(
            a.b.nbr, a.b.stride, i, a.b.cnt[i],
            [&](uint32_t j) { return Nbr3{j, ld4(a.pf, j), ld4(a.xs, j), ld4(v3b, j)}; },
            [&](const Nbr3& r) {
                const float3 pj = f3(r.a);
                const float3 d = pi - pj;
                const float W = poly6_d2(dot3(d, d), s);
                const float3 dv = inv_dt * (pj - f3(r.b)) - v2i;
                const float3 vbj = f3(r.c);
                v2b = v2b + (a.visc * W) * (vbj - vbi);
                cpart += W * dot3(vbi, dv);
                pacc = pacc + (a.visc * dot3(dv, vbi - vbj)) * poly6_grad(d, s);
            });
        const uint32_t gid = a.b.ids[i];
        const float3 g = inv_dt * v2b;
        float4 p = ld4(pb, gid);
        p.x += pacc.x + g.x;
        p.y += pacc.y + g.y;
        p.z += pacc.z + g.z;
        st4(pb, gid, p);
        st4(x0b, gid, make_float4(-g.x, -g.y, -g.z, 0.0f));
    }
    wave_add(acc + 0, cpart);
}

// ---------------------------------------------------------------------------
// Constraint iteration (one build)
// ---------------------------------------------------------------------------
struct IterArgs {
    pbfint::AdjBuild b;
    const pbf_float4* pk;    // positions entering the iteration, lambda in .w
    const pbf_float4* pre;   // positions after Delta-p, before the collision solve
    size_t n;
    float h, mass, rho0, eps, sk, sdq, radius;
    int sn;
    Colliders col;
};

// The collision solve's reverse (J^T of the replayed fixed point from the
// taped pre-projection position), then the lambda terms (density and
// gradient sums in FP64, as the forward) for the next two passes.
__global__ void __launch_bounds__(kBlock)
    k_adj_it1(IterArgs a, pbf_float4* __restrict__ pb, pbf_float4* __restrict__ prb,
              pbf_float4* __restrict__ gsum, float* __restrict__ cterm) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t gid = a.b.ids[i];
    const M3 J = collide_jacobian(ld4(a.pre, i), a.col, a.radius);
    const float3 rr = m3_tmul(J, f3(ld4(pb, gid)));
    const float4 r = make_float4(rr.x, rr.y, rr.z, 0.0f);
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
}

// Delta-p_i = (m / rho0) sum_j (l_i + l_j + s_ij) G_ij.  Pulled, with r the
// projected adjoint: l-bar_i = gs sum_j (r_i - r_j) . G_ij and d-bar_ij -
// d-bar_ji = gs coef_ij J (r_i - r_j) + gs ((r_i - r_j) . G_ij) grad s_ij.
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
    const float3 ri = f3(ld4(prb, i));
    float lb = 0.0f;
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) { return Nbr2{j, ld4(a.pk, j), ld4(prb, j)}; },
        [&](const Nbr2& r) {
            const float3 d = pi - f3(r.a);
            const float d2 = dot3(d, d);
            const float3 G = spiky(d, s);
            const float3 rr = ri - f3(r.b);
            const float q = gs * dot3(rr, G);
            lb += q;
            const float coef = pk.w + r.a.w + scorr_value(d2, s.h2, sc);
            pacc = pacc + (gs * coef) * spiky_jac(d, rr, s) + q * scorr_grad(d, s.h2, sc);
        });
    lamb[i] = lb;
    const uint32_t gid = a.b.ids[i];
    float4 p = ld4(pb, gid);
    p.x += pacc.x;
    p.y += pacc.y;
    p.z += pacc.z;
    st4(pb, gid, p);
}

// C-bar and S-bar of lambda_i = -C_i / (S_i + eps).
__device__ __forceinline__ float2 lambda_bars(float lb, float C, float S, float eps) {
    const float den = S + eps;
    return make_float2(-lb / den, lb * C / (den * den));
}

// lambda_i = -C_i / (S_i + eps), C_i = rho_i / rho0 - 1, S_i = sum_j |g_ij|^2 +
// |sum_j g_ij|^2, g_ij = -gs G_ij.  Pulled: d-bar_ij - d-bar_ji = gs (C-bar_i
// + C-bar_j) grad W_ij - gs J (gbar_ij - gbar_ji), gbar_ij = 2 S-bar_i (g_ij +
// gsum_i), gbar_ji = 2 S-bar_j (-g_ij + gsum_j).
__global__ void __launch_bounds__(kBlock)
    k_adj_it3(IterArgs a, const pbf_float4* __restrict__ gsum, const float* __restrict__ cterm,
              const float* __restrict__ lamb, pbf_float4* __restrict__ pb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Sph s = sph_consts(a.h);
    const float gs = a.mass / a.rho0;
    const float4 gsi = ld4(gsum, i);
    const float2 bi = lambda_bars(lamb[i], cterm[i], gsi.w, a.eps);
    const float3 pi = f3(ld4(a.pk, i));
    const float3 gsv = f3(gsi);
    float3 pacc = make_float3(0.0f, 0.0f, 0.0f);
    for_each_neighbour<true>(
        a.b.nbr, a.b.stride, i, a.b.cnt[i],
        [&](uint32_t j) {
            return Nbr3{j, ld4(a.pk, j), ld4(gsum, j), make_float4(lamb[j], cterm[j], 0.0f, 0.0f)};
        },
        [&](const Nbr3& r) {
            const float2 bj = lambda_bars(r.c.x, r.c.y, r.b.w, a.eps);
            const float3 d = pi - f3(r.a);
            const float3 g = (-gs) * spiky(d, s);
            const float3 gbi = (2.0f * bi.y) * (g + gsv);
            const float3 gbj = (2.0f * bj.y) * (f3(r.b) - g);
            pacc = pacc + ((bi.x + bj.x) * gs) * poly6_grad(d, s) +
                   (-gs) * spiky_jac(d, gbi - gbj, s);
        });
    const uint32_t gid = a.b.ids[i];
    float4 p = ld4(pb, gid);
    p.x += pacc.x;
    p.y += pacc.y;
    p.z += pacc.z;
    st4(pb, gid, p);
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

// One grid build into the tape: the list groups that hold entries, the
// counts and the slot-to-id map (tape_build).
__global__ void __launch_bounds__(kBlock)
    k_tape_build(const uint4* __restrict__ src, uint4* __restrict__ dst,
                 const int* __restrict__ cnt, int* __restrict__ cntOut,
                 const uint32_t* __restrict__ ids, uint32_t* __restrict__ idsOut, size_t n,
                 size_t stride, int maxGroups) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i];
    cntOut[i] = c;
    idsOut[i] = ids[i];
    const int groups = min((c + 3) / 4, maxGroups);
    for (int g = 0; g < groups; ++g) dst[static_cast<size_t>(g) * stride + i] = src[static_cast<size_t>(g) * stride + i];
}

// Three record arrays into the tape (tape_records).
__global__ void __launch_bounds__(kBlock)
    k_tape_records(const float4* __restrict__ a, float4* __restrict__ aOut,
                   const float4* __restrict__ b, float4* __restrict__ bOut,
                   const float4* __restrict__ c, float4* __restrict__ cOut, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aOut[i] = a[i];
    bOut[i] = b[i];
    cOut[i] = c[i];
}

}  // namespace

namespace pbfint {

pbf_status tape_build(const uint32_t* lists, uint32_t* listsOut, const int* cnt, int* cntOut,
                      const uint32_t* ids, uint32_t* idsOut, size_t n, size_t stride,
                      int maxGroups, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_tape_build, grid_for(n), kBlock, stream, reinterpret_cast<const uint4*>(lists),
               reinterpret_cast<uint4*>(listsOut), cnt, cntOut, ids, idsOut, n, stride, maxGroups);
}

pbf_status tape_records(const pbf_float4* a, pbf_float4* aOut, const pbf_float4* b,
                        pbf_float4* bOut, const pbf_float4* c, pbf_float4* cOut, size_t n,
                        void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_tape_records, grid_for(n), kBlock, stream, reinterpret_cast<const float4*>(a),
               reinterpret_cast<float4*>(aOut), reinterpret_cast<const float4*>(b),
               reinterpret_cast<float4*>(bOut), reinterpret_cast<const float4*>(c),
               reinterpret_cast<float4*>(cOut), n);
}

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
    a.col = prm.col;
    const unsigned g = grid_for(n);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    pbf_status rc;
    hipLaunchKernelGGL(k_adj_post1, dim3(g), dim3(kBlock), 0, st, a, w.vb, w.v1b, w.v3b, w.etab,
                       w.wb, w.pb, w.acc);
    if ((rc = pbf_hip_check(hipGetLastError())) != PBF_OK) return rc;
    hipLaunchKernelGGL(k_adj_post2, dim3(g), dim3(kBlock), 0, st, a, w.etab, w.wb, w.pb);
    if ((rc = pbf_hip_check(hipGetLastError())) != PBF_OK) return rc;
    hipLaunchKernelGGL(k_adj_post4, dim3(g), dim3(kBlock), 0, st, a, w.wb, w.v3b, w.pb);
    if ((rc = pbf_hip_check(hipGetLastError())) != PBF_OK) return rc;
    PBF_LAUNCH(k_adj_post5, g, kBlock, stream, a, w.v3b, w.pb, w.x0b, w.acc);
}

pbf_status adjoint_iteration(const AdjIter& t, const AdjBuffers& w, size_t n, const AdjParams& prm,
                             void* stream) {
    if (n == 0) return PBF_OK;
    IterArgs a;
    a.b = t.build;
    a.pk = t.pk;
    a.pre = t.pre;
    a.n = n;
    a.h = prm.h;
    a.mass = prm.mass;
    a.rho0 = prm.rho0;
    a.eps = prm.lambdaEps;
    a.sk = prm.scorrK;
    a.sn = prm.scorrN;
    a.sdq = prm.scorrDq;
    a.radius = prm.radius;
    a.col = prm.col;
    const unsigned g = grid_for(n);
    pbf_status st;
    hipLaunchKernelGGL(k_adj_it1, dim3(g), dim3(kBlock), 0, static_cast<hipStream_t>(stream), a,
                       w.pb, w.prb, w.gsum, w.cterm);
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
