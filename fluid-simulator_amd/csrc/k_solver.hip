// PBF particle kernels for gfx950: the reference's free kernels (same
// argument meaning, slot-major neighbour lists) and the fused kernels the
// solver step launches.
//
// Reference: src/pbf/integration.cu, src/pbf/pbf_solver_kernels.cu,
// include/pbf/pbf_solver_kernels.cuh, src/pbf/collision_kernels.cu.
#include "pbf_device.hpp"
#include "pbf_internal.hpp"
#include "pbf_launch.hpp"

using namespace pbfdev;

namespace {

// The step's list kernels run the kFast form: the hardware reciprocal square
// root in the spiky gradient (spiky_grad_rsq) and workgroups laid out
// contiguously per XCD.  The C ABI launchers run the reference-shaped form
// (correctly rounded sqrt and division, plain block order).
template <bool kFast>
__device__ __forceinline__ float3 grad(float dx, float dy, float dz, float d2, const Sph& s) {
    return kFast ? spiky_grad_rsq(dx, dy, dz, d2, s) : spiky_grad(dx, dy, dz, d2, s);
}

template <bool kFast>
__device__ __forceinline__ size_t thread_index() {
    const unsigned b = kFast ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    return static_cast<size_t>(b) * blockDim.x + threadIdx.x;
}

// Particles per thread (a launch tuning): workgroup b takes the ppt blocks
// of blockDim particles from b * ppt on, one after another, so the lanes of
// a wave still touch consecutive particles at every step.  body(i) returns
// early for particles it skips.
template <bool kFast, class Body>
__device__ __forceinline__ void for_each_particle(int ppt, Body body) {
    const size_t b = kFast ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    for (int q = 0; q < ppt; ++q)
        body((b * static_cast<size_t>(ppt) + static_cast<size_t>(q)) * blockDim.x + threadIdx.x);
}

// Cell key of position p: the scene's own key, offset into the global key
// range in a batched step (the scene of slot i).
template <bool kBatched>
__device__ __forceinline__ uint32_t cell_key(float4 p, size_t i, pbf_float3 mn, float cs,
                                             pbf_int3 gs, const pbfint::SceneTable& st) {
    if (!kBatched) return hash_key(p, mn, cs, gs);
    const pbfint::SceneGeom& sg = st.geom[st.sceneOf[i]];
    return sg.cellOff + hash_key(p, sg.minBounds, cs, sg.dims);
}

// cudaMemset(-1) of the next grid build's cellStart (spatial_grid.cu:203-204),
// folded into a kernel that runs between the previous build's list sweep
// and the next build's ranges: thread t of the launch writes cells t,
// t + threads, ...  Only cellStart is cleared: the list build reads a
// cell's end only where its start is set.
__device__ __forceinline__ void clear_cells(int* __restrict__ cells, size_t count) {
    if (cells == nullptr) return;
    const size_t threads = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t c = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < count;
         c += threads)
        cells[c] = -1;
}

// The colliders of slot i: in a batched step, its scene's container.
template <bool kBatched>
__device__ __forceinline__ Colliders colliders_of(const Colliders& col, size_t i,
                                                  const pbfint::SceneTable& st) {
    Colliders c = col;
    if (kBatched) c.container = st.containers + st.sceneOf[i];
    return c;
}

struct PosLambda {
    float4 p;
    float l;
};
struct PosPair {
    float4 a, b;
};

// ---------------------------------------------------------------------------
// Reference-shaped kernels
// ---------------------------------------------------------------------------

// integrate (src/pbf/integration.cu:6-32).
__global__ void __launch_bounds__(kBlock) k_integrate(const pbf_float4* __restrict__ x,
                                                      pbf_float4* __restrict__ p,
                                                      pbf_float4* __restrict__ v, int n, float dt,
                                                      pbf_float3 g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 vi = ld4(v, i);
    const float4 xi = ld4(x, i);
    vi.x = vi.x + g.x * dt;
    vi.y = vi.y + g.y * dt;
    vi.z = vi.z + g.z * dt;
    float4 pi;
    pi.x = xi.x + vi.x * dt;
    pi.y = xi.y + vi.y * dt;
    pi.z = xi.z + vi.z * dt;
    pi.w = xi.w;
    st4(v, i, vi);
    st4(p, i, pi);
}

// computeLambda (src/pbf/pbf_solver_kernels.cu:26-105).  lambda_i is stored
// at lambda[i * lstride]: the C ABI passes its own array (lstride 1), the
// step passes the .w words of p itself (lstride 4), so that Delta-p reads a
// neighbour's position and lambda in one 16-byte gather.  Storing .w while
// other lanes read .xyz of the same records touches disjoint bytes, hence
// neither pointer is restrict-qualified.
// kSkin: the list is a skin list; its entries are tested against h (the
// reference's inclusive FP32 test), and more than exactMax within h raise
// the overflow bit, as the grid build the skin list stands in for would.
template <bool kFast, bool kSkin = false>
__global__ void __launch_bounds__(kBlock)
    k_lambda_list(const pbf_float4* p, const uint32_t* __restrict__ nbr,
                  const int* __restrict__ cnt, int maxN, size_t n, float h, float mass, float rho0,
                  float eps, float* lambda, size_t lstride, size_t stride,
                  uint32_t* __restrict__ status, int exactMax, int ppt,
                  pbf_float4* __restrict__ tapeOut) {
    for_each_particle<kFast>(ppt, [&](const size_t i) {
    if (i >= n) return;
    if (status != nullptr && *status != 0u) return;
    const float4 pi = ld4(p, i);
    const int c = cnt[i];
    if (!valid_count(c, maxN)) {
        lambda[i * lstride] = qnan();
        if (tapeOut != nullptr) st4(tapeOut, i, make_float4(pi.x, pi.y, pi.z, qnan()));
        return;
    }
    const Sph s = sph_consts(h);
    const float gs = mass / rho0;
    // Per-neighbour terms in FP32, their sums in FP64: the density sits
    // close to the rest density, so C = rho / rho0 - 1 cancels most of an
    // FP32 sum's significant bits.
    double rho = static_cast<double>(mass * s.poly_norm);
    double gix = 0.0, giy = 0.0, giz = 0.0, sg = 0.0;
    int within = 0;
    for_each_neighbour<kFast, kFast ? 2 : 1>(
        nbr, stride, i, c, [&](uint32_t j) { return ld4(p, j); },
        [&](const float4& pj) {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            float d2;
            if (kSkin) {
                d2 = dist2_exact(dx, dy, dz);
                within += d2 <= s.h2 ? 1 : 0;
            } else {
                d2 = dist2_fma(dx, dy, dz);
            }
            rho += static_cast<double>(mass * poly6_d2(d2, s));
            const float3 gw = grad<kFast>(dx, dy, dz, d2, s);
            const float gjx = -gs * gw.x, gjy = -gs * gw.y, gjz = -gs * gw.z;
            sg += static_cast<double>(dist2_fma(gjx, gjy, gjz));
            gix -= static_cast<double>(gjx);
            giy -= static_cast<double>(gjy);
            giz -= static_cast<double>(gjz);
        });
    sg += dist2_fma(gix, giy, giz);
    const double C = rho / static_cast<double>(rho0) - 1.0;
    const float lam = static_cast<float>(-C / (sg + static_cast<double>(eps)));
    lambda[i * lstride] = lam;
    if (tapeOut != nullptr) st4(tapeOut, i, make_float4(pi.x, pi.y, pi.z, lam));
    if (kSkin && within > exactMax) atomicOr(status, PBF_STATUS_NEIGHBOR_OVERFLOW);
    });
}

// computeDeltaPosition (src/pbf/pbf_solver_kernels.cu:108-182).
__global__ void __launch_bounds__(kBlock)
    k_delta_list(const pbf_float4* __restrict__ p, const uint32_t* __restrict__ nbr,
                 const int* __restrict__ cnt, int maxN, const float* __restrict__ lambda, size_t n,
                 float h, float mass, float rho0, float sk, int sn, float sdq,
                 pbf_float4* __restrict__ delta, size_t stride) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pi = ld4(p, i);
    const float li = lambda[i];
    const int c = cnt[i];
    if (!valid_count(c, maxN)) {
        st4(delta, i, make_float4(qnan(), qnan(), qnan(), 0.0f));
        return;
    }
    const Sph s = sph_consts(h);
    const Scorr sc = scorr_consts(h, sk, sn, sdq);
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int q = 0; q < c; ++q) {
        const uint32_t j = nbr[static_cast<size_t>(q) * stride + i];
        const float4 pj = ld4(p, j);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float d2 = dist2_fma(dx, dy, dz);
        const float3 gw = spiky_grad(dx, dy, dz, d2, s);
        const float ls = li + lambda[j] + scorr_value(d2, s.h2, sc);
        ax += ls * gw.x;
        ay += ls * gw.y;
        az += ls * gw.z;
    }
    const float cs = mass / rho0;
    st4(delta, i, make_float4(cs * ax, cs * ay, cs * az, 0.0f));
}

// applyDeltaPosition (src/pbf/pbf_solver_kernels.cu:184-193).
__global__ void __launch_bounds__(kBlock) k_apply_delta(pbf_float4* __restrict__ p,
                                                        const pbf_float4* __restrict__ d,
                                                        size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    const float4 di = ld4(d, i);
    pi.x += di.x;
    pi.y += di.y;
    pi.z += di.z;
    st4(p, i, pi);
}

// updateVelocityAndPosition (src/pbf/pbf_solver_kernels.cu:195-211).
__global__ void __launch_bounds__(kBlock) k_update_velocity(pbf_float4* __restrict__ x,
                                                            const pbf_float4* __restrict__ p,
                                                            pbf_float4* __restrict__ v, size_t n,
                                                            float inv_dt) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = ld4(x, i);
    const float4 pi = ld4(p, i);
    float4 vi = ld4(v, i);
    vi.x = (pi.x - xi.x) * inv_dt;
    vi.y = (pi.y - xi.y) * inv_dt;
    vi.z = (pi.z - xi.z) * inv_dt;
    st4(v, i, vi);
    st4(x, i, pi);
}

// applyXsphViscosity (src/pbf/pbf_solver_kernels.cu:213-266).
__global__ void __launch_bounds__(kBlock)
    k_xsph_list(const pbf_float4* __restrict__ p, const uint32_t* __restrict__ nbr,
                const int* __restrict__ cnt, int maxN, const pbf_float4* __restrict__ vin,
                pbf_float4* __restrict__ vout, size_t n, float h, float visc, size_t stride) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 vi = ld4(vin, i);
    const int c = cnt[i];
    if (!valid_count(c, maxN)) {
        st4(vout, i, make_float4(qnan(), qnan(), qnan(), vi.w));
        return;
    }
    if (visc == 0.0f) {
        st4(vout, i, vi);
        return;
    }
    const Sph s = sph_consts(h);
    const float4 pi = ld4(p, i);
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (int q = 0; q < c; ++q) {
        const uint32_t j = nbr[static_cast<size_t>(q) * stride + i];
        const float4 pj = ld4(p, j);
        const float4 vj = ld4(vin, j);
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float w = poly6_d2(dist2_fma(dx, dy, dz), s);
        cx += (vj.x - vi.x) * w;
        cy += (vj.y - vi.y) * w;
        cz += (vj.z - vi.z) * w;
    }
    st4(vout, i, make_float4(vi.x + visc * cx, vi.y + visc * cy, vi.z + visc * cz, vi.w));
}

// computeVorticity (src/pbf/pbf_solver_kernels.cu:268-315).  The kFast form
// also stores |omega| in .w (the reference writes 0 there) and the record
// (p.xyz, |omega|) in pw, so that the step's confinement kernel reads a
// neighbour's position and |omega| in one 16-byte gather.
template <bool kFast>
__global__ void __launch_bounds__(kBlock)
    k_vorticity_list(const pbf_float4* __restrict__ p, const pbf_float4* __restrict__ v,
                     const uint32_t* __restrict__ nbr, const int* __restrict__ cnt, int maxN,
                     size_t n, float h, pbf_float4* __restrict__ omega,
                     pbf_float4* __restrict__ pw, size_t stride, uint32_t* __restrict__ status,
                     int exactMax, int ppt, pbf_float4* __restrict__ tapeOut) {
    for_each_particle<kFast>(ppt, [&](const size_t i) {
    if (i >= n) return;
    const int c = cnt[i];
    const float4 pi = ld4(p, i);
    if (!valid_count(c, maxN)) {
        st4(omega, i, make_float4(qnan(), qnan(), qnan(), 0.0f));
        if (tapeOut != nullptr) st4(tapeOut, i, make_float4(qnan(), qnan(), qnan(), 0.0f));
        if (kFast) st4(pw, i, make_float4(pi.x, pi.y, pi.z, qnan()));
        return;
    }
    const Sph s = sph_consts(h);
    const float4 vi = ld4(v, i);
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    int within = 0;
    for_each_neighbour<kFast, kFast ? 2 : 1>(
        nbr, stride, i, c, [&](uint32_t j) { return PosPair{ld4(p, j), ld4(v, j)}; },
        [&](const PosPair& r) {
            const float4 &pj = r.a, &vj = r.b;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            if (exactMax > 0) within += dist2_exact(dx, dy, dz) <= s.h2 ? 1 : 0;
            const float3 g = grad<kFast>(dx, dy, dz, dist2_fma(dx, dy, dz), s);
            const float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
            ox += cross_fma(g.y, uz, g.z, uy);
            oy += cross_fma(g.z, ux, g.x, uz);
            oz += cross_fma(g.x, uy, g.y, ux);
        });
    if (exactMax > 0 && within > exactMax) atomicOr(status, PBF_STATUS_NEIGHBOR_OVERFLOW);
    const float wl = kFast ? __builtin_sqrtf(dist2_fma(ox, oy, oz)) : 0.0f;
    st4(omega, i, make_float4(ox, oy, oz, wl));
    if (tapeOut != nullptr) st4(tapeOut, i, make_float4(ox, oy, oz, wl));
    if (kFast) st4(pw, i, make_float4(pi.x, pi.y, pi.z, wl));
    });
}

// applyVorticityConfinement (src/pbf/pbf_solver_kernels.cu:317-385) as a
// device function, shared by the API kernel and the fused confinement +
// contact kernel.  The kFast form reads p = the (position, |omega|) records
// of the kFast vorticity kernel: one gather per neighbour.
template <bool kFast>
__device__ __forceinline__ void confine(size_t i, const pbf_float4* __restrict__ p,
                                        const uint32_t* __restrict__ nbr, int c,
                                        const pbf_float4* __restrict__ omega, float4 vi,
                                        const Sph& s, float dt, float strength, size_t stride,
                                        float4& out) {
    const float4 pi = ld4(p, i);
    const float4 wi = ld4(omega, i);
    const float wil = kFast ? pi.w : __builtin_sqrtf(dist2_fma(wi.x, wi.y, wi.z));
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    for_each_neighbour<kFast, kFast ? 2 : 1>(
        nbr, stride, i, c,
        [&](uint32_t j) {
            const float4 pj = ld4(p, j);
            float wjl;
            if (kFast) {
                wjl = pj.w;
            } else {
                const float4 wj = ld4(omega, j);
                wjl = __builtin_sqrtf(dist2_fma(wj.x, wj.y, wj.z));
            }
            return PosLambda{pj, wjl};
        },
        [&](const PosLambda& r) {
            const float4& pj = r.p;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float3 g = grad<kFast>(dx, dy, dz, dist2_fma(dx, dy, dz), s);
            const float md = r.l - wil;
            ex += md * g.x;
            ey += md * g.y;
            ez += md * g.z;
        });
    const float el = __builtin_sqrtf(dist2_fma(ex, ey, ez));
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (el > 1.0e-6f) {  // normalizationEpsilon, pbf_solver_kernels.cu:367
        nx = ex / el;
        ny = ey / el;
        nz = ez / el;
    }
    const float fx = strength * cross_fma(ny, wi.z, nz, wi.y);
    const float fy = strength * cross_fma(nz, wi.x, nx, wi.z);
    const float fz = strength * cross_fma(nx, wi.y, ny, wi.x);
    out = make_float4(vi.x + dt * fx, vi.y + dt * fy, vi.z + dt * fz, vi.w);
}

__global__ void __launch_bounds__(kBlock)
    k_confine_list(const pbf_float4* __restrict__ p, const uint32_t* __restrict__ nbr,
                   const int* __restrict__ cnt, int maxN, const pbf_float4* __restrict__ omega,
                   const pbf_float4* __restrict__ vin, pbf_float4* __restrict__ vout, size_t n,
                   float h, float dt, float strength, size_t stride) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 vi = ld4(vin, i);
    const int c = cnt[i];
    if (!valid_count(c, maxN)) {
        st4(vout, i, make_float4(qnan(), qnan(), qnan(), vi.w));
        return;
    }
    const Sph s = sph_consts(h);
    float4 out;
    confine<false>(i, p, nbr, c, omega, vi, s, dt, strength, stride, out);
    st4(vout, i, out);
}

// gatherSolverState (src/pbf/pbf_solver_kernels.cu:387-407).
__global__ void __launch_bounds__(kBlock)
    k_gather(const uint32_t* __restrict__ idx, const pbf_float4* __restrict__ x,
             const pbf_float4* __restrict__ p, const pbf_float4* __restrict__ v,
             const pbf_float4* __restrict__ vin, const uint32_t* __restrict__ id,
             pbf_float4* __restrict__ xo, pbf_float4* __restrict__ po, pbf_float4* __restrict__ vo,
             pbf_float4* __restrict__ vino, uint32_t* __restrict__ ido, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = idx[i];
    st4(xo, i, ld4(x, s));
    st4(po, i, ld4(p, s));
    st4(vo, i, ld4(v, s));
    st4(vino, i, ld4(vin, s));
    ido[i] = id[s];
}

// initializeStableParticleIds (src/pbf/pbf_solver_kernels.cu:409-416).
__global__ void __launch_bounds__(kBlock) k_init_ids(uint32_t* __restrict__ id, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) id[i] = static_cast<uint32_t>(i);
}

// scatterFloat4ByStableId (src/pbf/pbf_solver_kernels.cu:418-427).
__global__ void __launch_bounds__(kBlock) k_scatter_by_id(const pbf_float4* __restrict__ vals,
                                                          const uint32_t* __restrict__ id,
                                                          pbf_float4* __restrict__ out, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) st4(out, id[i], ld4(vals, i));
}

// Collision kernels (src/pbf/collision_kernels.cu:47-234): one projection
// per launch, as the reference's CollisionSystem passes launch them.
__global__ void __launch_bounds__(kBlock) k_solve_container(pbf_float4* __restrict__ p, size_t n,
                                                            const pbf_container* __restrict__ c,
                                                            float r, int* __restrict__ flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    const pbf_container k = *c;
    if (project_container(pi, k, r) && flag != nullptr) atomicOr(flag, 1);
    st4(p, i, pi);
}

__global__ void __launch_bounds__(kBlock) k_solve_spheres(pbf_float4* __restrict__ p, size_t n,
                                                          const pbf_sphere* __restrict__ s,
                                                          size_t count, float r,
                                                          int* __restrict__ flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    bool marked = false;
    for (size_t k = 0; k < count; ++k) marked |= project_sphere(pi, s[k], r);
    if (marked && flag != nullptr) atomicOr(flag, 1);
    st4(p, i, pi);
}

__global__ void __launch_bounds__(kBlock) k_solve_boxes(pbf_float4* __restrict__ p, size_t n,
                                                        const pbf_box* __restrict__ b,
                                                        size_t count, float r,
                                                        int* __restrict__ flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    bool marked = false;
    for (size_t k = 0; k < count; ++k) marked |= project_box(pi, b[k], r);
    if (marked && flag != nullptr) atomicOr(flag, 1);
    st4(p, i, pi);
}

__global__ void __launch_bounds__(kBlock) k_solve_planes(pbf_float4* __restrict__ p, size_t n,
                                                         const pbf_plane* __restrict__ pl,
                                                         size_t count, float r,
                                                         int* __restrict__ flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    bool marked = false;
    for (size_t k = 0; k < count; ++k) marked |= project_plane(pi, pl[k], r);
    if (marked && flag != nullptr) atomicOr(flag, 1);
    st4(p, i, pi);
}

// resolveVelocitiesKernel (src/pbf/collision_kernels.cu:236-371).
__global__ void __launch_bounds__(kBlock)
    k_resolve_velocities(const pbf_float4* __restrict__ x, const pbf_float4* vin,
                         pbf_float4* v, size_t n, Colliders col, float r, float e, float f) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = ld4(x, i);
    const float4 vini = ld4(vin, i);
    float4 vi = ld4(v, i);
    resolve_velocity(xi, vini, vi, col, r, e, f);
    st4(v, i, vi);
}

// Batched device evaluation of the inline helpers.
__global__ void k_poly6(const pbf_float3* __restrict__ d, size_t n, float h, float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Sph s = sph_consts(h);
    const pbf_float3 q = d[i];
    out[i] = poly6_d2(dist2_fma(q.x, q.y, q.z), s);
}

__global__ void k_spiky(const pbf_float3* __restrict__ d, size_t n, float h,
                        pbf_float3* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Sph s = sph_consts(h);
    const pbf_float3 q = d[i];
    const float3 g = spiky_grad(q.x, q.y, q.z, dist2_fma(q.x, q.y, q.z), s);
    out[i] = pbf_float3{g.x, g.y, g.z};
}

__global__ void k_scorr(const float* __restrict__ d2, size_t n, float h, float k, int sn, float dq,
                        float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Scorr c = scorr_consts(h, k, sn, dq);
    out[i] = scorr_value(d2[i], h * h, c);
}

// ---------------------------------------------------------------------------
// Fused solver kernels
// ---------------------------------------------------------------------------

// integrate + collisionInputVelocities snapshot + cell key of p*
// (pbf_solver.cu:199-208 and spatial_grid.cu:100-120 in one pass).
// With box set (skin lists), the key is that of p*'s container point.
template <bool kBatched>
__global__ void __launch_bounds__(kBlock)
    k_predict_key(const pbf_float4* __restrict__ x, pbf_float4* __restrict__ p,
                  pbf_float4* __restrict__ v, pbf_float4* __restrict__ vin,
                  uint32_t* __restrict__ key, int n, float dt, pbf_float3 g, pbf_float3 mn,
                  float cs, pbf_int3 gs, const pbf_container* __restrict__ box, float radius,
                  pbfint::SceneTable st, int* __restrict__ clearCells, size_t clearCount,
                  const uint32_t* __restrict__ status) {
    clear_cells(clearCells, clearCount);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (*status != 0u) return;
    float4 vi = ld4(v, i);
    const float4 xi = ld4(x, i);
    vi.x = vi.x + g.x * dt;
    vi.y = vi.y + g.y * dt;
    vi.z = vi.z + g.z * dt;
    float4 pi;
    pi.x = xi.x + vi.x * dt;
    pi.y = xi.y + vi.y * dt;
    pi.z = xi.z + vi.z * dt;
    pi.w = xi.w;
    st4(v, i, vi);
    st4(vin, i, vi);
    st4(p, i, pi);
    float4 kp = pi;
    if (box != nullptr) {
        const pbf_container& c = kBatched ? st.containers[st.sceneOf[i]] : *box;
        const float3 q = container_point(pi, c, radius);
        kp = make_float4(q.x, q.y, q.z, pi.w);
    }
    key[i] = cell_key<kBatched>(kp, static_cast<size_t>(i), mn, cs, gs, st);
}

// The frame checkpoint of the skin lists: the state a stale frame restarts from.
__global__ void __launch_bounds__(kBlock)
    k_checkpoint(const pbf_float4* __restrict__ x, const pbf_float4* __restrict__ v,
                 const uint32_t* __restrict__ id, pbf_float4* __restrict__ cx,
                 pbf_float4* __restrict__ cv, uint32_t* __restrict__ cid, size_t n,
                 const uint32_t* __restrict__ status) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (*status != 0u) return;
    st4(cx, i, ld4(x, i));
    st4(cv, i, ld4(v, i));
    cid[i] = id[i];
}

__global__ void k_frame_done(uint32_t* __restrict__ status) {
    if (threadIdx.x == 0 && status[0] == 0u) status[1] += 1u;
}

// Cell keys of the current working set (computeParticleKeys,
// spatial_grid.cu:100-120); the slot index is implicit.
template <bool kBatched>
__global__ void __launch_bounds__(kBlock)
    k_keys(const pbf_float4* __restrict__ p, uint32_t* __restrict__ key, size_t n, pbf_float3 mn,
           float cs, pbf_int3 gs, pbfint::SceneTable st) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = cell_key<kBatched>(ld4(p, i), i, mn, cs, gs, st);
}

// computeDeltaPosition + applyDeltaPosition + CollisionSystem::solve for one
// constraint iteration (pbf_solver.cu:240-257).  Reads p* from pin and writes
// the corrected, projected p* to pout (Jacobi), so every thread sees the
// positions the reference's separate Delta-p kernel sees.  pin[].w holds
// lambda (stored there by the lambda pass); pout[].w gets back the
// particle's own .w, which x carries.
// kSkin: the list is a skin list (out-of-range entries add exact zeros),
// and a particle that ends farther than sqrt(stale2) from the container
// point of its p0 (the position the lists were built around) marks the
// lists stale.
template <bool kBatched, bool kSkin>
__global__ void __launch_bounds__(kBlock)
    k_delta_apply_collide(const pbf_float4* __restrict__ pin, pbf_float4* __restrict__ pout,
                          const uint32_t* __restrict__ nbr, const int* __restrict__ cnt,
                          int maxN, const pbf_float4* __restrict__ x, size_t n, float h,
                          float mass, float rho0, float sk, int sn, float sdq, size_t stride,
                          Colliders col, float radius, pbfint::SceneTable st,
                          const pbf_float4* p0, float stale2, uint32_t* __restrict__ key,
                          pbf_float3 kmn, float kcs, pbf_int3 kgs, int* __restrict__ clearCells,
                          size_t clearCount, const uint32_t* __restrict__ xperm,
                          pbf_float4* __restrict__ prePos, uint32_t* __restrict__ status, int ppt) {
    clear_cells(clearCells, clearCount);
    for_each_particle<true>(ppt, [&](const size_t i) {
    if (i >= n) return;
    if (*status != 0u) return;
    float4 pi = ld4(pin, i);
    const int c = cnt[i];
    float dpx, dpy, dpz;
    if (!valid_count(c, maxN)) {
        dpx = dpy = dpz = qnan();
    } else {
        const Sph s = sph_consts(h);
        const Scorr sc = scorr_consts(h, sk, sn, sdq);
        const float li = pi.w;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for_each_neighbour<true, 2>(
            nbr, stride, i, c, [&](uint32_t j) { return ld4(pin, j); },
            [&](const float4& pj) {
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float d2 = dist2_fma(dx, dy, dz);
                const float3 gw = grad<true>(dx, dy, dz, d2, s);
                const float ls = li + pj.w + scorr_value(d2, s.h2, sc);
                ax += ls * gw.x;
                ay += ls * gw.y;
                az += ls * gw.z;
            });
        const float cs = mass / rho0;
        dpx = cs * ax;
        dpy = cs * ay;
        dpz = cs * az;
    }
    pi.x += dpx;
    pi.y += dpy;
    pi.z += dpz;
    pi.w = x[xperm != nullptr ? xperm[i] : i].w;
    if (prePos != nullptr) st4(prePos, i, pi);
    const Colliders ci = colliders_of<kBatched>(col, i, st);
    if (!collision_fixed_point(pi, ci, radius))
        atomicOr(status, PBF_STATUS_COLLISION_NONCONVERGENCE);
    st4(pout, i, pi);
    if (key != nullptr) key[i] = cell_key<kBatched>(pi, i, kmn, kcs, kgs, st);
    if (kSkin) {
        const float3 q = container_point(ld4(p0, i), *ci.container, radius);
        const float ex = pi.x - q.x, ey = pi.y - q.y, ez = pi.z - q.z;
        if (!(dist2_fma(ex, ey, ez) <= stale2)) atomicOr(status, PBF_STATUS_LIST_STALE);
    }
    });
}

// updateVelocityAndPosition + applyXsphViscosity (pbf_solver.cu:284-298).
// Velocities v = (p* - x) / dt are rebuilt on the fly for i and every
// neighbour with the same FP32 expression the reference stores, so the
// stored intermediate velocity buffer is not needed.  x <- p* is a buffer
// swap on the host.
// kSkin: skin lists; more than exactMax neighbours within h raise the
// overflow bit (the post-solve build's check).
template <bool kSkin>
__global__ void __launch_bounds__(kBlock)
    k_update_xsph(const pbf_float4* __restrict__ x, const pbf_float4* __restrict__ p,
                  const pbf_float4* __restrict__ vold, pbf_float4* __restrict__ vout,
                  const uint32_t* __restrict__ nbr, const int* __restrict__ cnt, int maxN,
                  size_t n, float h, float visc, float inv_dt, size_t stride,
                  uint32_t* __restrict__ status, int exactMax, int ppt,
                  pbf_float4* __restrict__ tapeOut) {
    for_each_particle<true>(ppt, [&](const size_t i) {
    if (i >= n) return;
    if (*status != 0u) return;
    const float4 xi = ld4(x, i);
    const float4 pi = ld4(p, i);
    const float vw = vold[i].w;
    const float vix = stored_velocity(pi.x, xi.x, inv_dt);
    const float viy = stored_velocity(pi.y, xi.y, inv_dt);
    const float viz = stored_velocity(pi.z, xi.z, inv_dt);
    const int c = cnt[i];
    if (!valid_count(c, maxN)) {
        st4(vout, i, make_float4(qnan(), qnan(), qnan(), vw));
        if (tapeOut != nullptr) st4(tapeOut, i, make_float4(qnan(), qnan(), qnan(), vw));
        return;
    }
    const Sph s = sph_consts(h);
    const float3 pi3 = make_float3(pi.x, pi.y, pi.z);
    const float3 vi3 = make_float3(vix, viy, viz);
    float3 cs = make_float3(0.0f, 0.0f, 0.0f);
    int within = 0;
    for_each_neighbour<true, 2>(
        nbr, stride, i, c, [&](uint32_t j) { return PosPair{ld4(p, j), ld4(x, j)}; },
        [&](const PosPair& r) {
            const float4 &pj = r.a, &xj = r.b;
            const float3 vj = make_float3(stored_velocity(pj.x, xj.x, inv_dt),
                                          stored_velocity(pj.y, xj.y, inv_dt),
                                          stored_velocity(pj.z, xj.z, inv_dt));
            if (kSkin) within += dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) <= s.h2;
            xsph_term(pi3, vi3, make_float3(pj.x, pj.y, pj.z), vj, s, cs);
        });
    const float4 vnew = make_float4(vix + visc * cs.x, viy + visc * cs.y, viz + visc * cs.z, vw);
    st4(vout, i, vnew);
    if (tapeOut != nullptr) st4(tapeOut, i, vnew);
    if (kSkin && within > exactMax) atomicOr(status, PBF_STATUS_NEIGHBOR_OVERFLOW);
    });
}

// updateVelocityAndPosition without XSPH, velocity only (x <- p* is a swap).
__global__ void __launch_bounds__(kBlock)
    k_update_velocity_only(const pbf_float4* __restrict__ x, const pbf_float4* __restrict__ p,
                           pbf_float4* __restrict__ v, size_t n, float inv_dt,
                           const uint32_t* __restrict__ status) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (*status != 0u) return;
    const float4 xi = ld4(x, i);
    const float4 pi = ld4(p, i);
    float4 vi = ld4(v, i);
    vi.x = (pi.x - xi.x) * inv_dt;
    vi.y = (pi.y - xi.y) * inv_dt;
    vi.z = (pi.z - xi.z) * inv_dt;
    st4(v, i, vi);
}

// applyVorticityConfinement + resolveVelocitiesKernel (pbf_solver.cu:309-323).
// pw holds the vorticity pass's (position, |omega|) records.
template <bool kBatched>
__global__ void __launch_bounds__(kBlock)
    k_confine_resolve(const pbf_float4* __restrict__ x, const pbf_float4* __restrict__ pw,
                      const uint32_t* __restrict__ nbr,
                      const int* __restrict__ cnt, int maxN, const pbf_float4* __restrict__ omega,
                      const pbf_float4* __restrict__ vcur, const pbf_float4* __restrict__ vsnap,
                      pbf_float4* __restrict__ vout, size_t n, float h, float dt, float strength,
                      size_t stride, Colliders col, float radius, float e, float f,
                      pbfint::SceneTable st, const uint32_t* __restrict__ status, int ppt) {
    for_each_particle<true>(ppt, [&](const size_t i) {
    if (i >= n) return;
    if (*status != 0u) return;
    const float4 vi = ld4(vcur, i);
    const int c = cnt[i];
    float4 out;
    if (!valid_count(c, maxN)) {
        out = make_float4(qnan(), qnan(), qnan(), vi.w);
    } else {
        const Sph s = sph_consts(h);
        confine<true>(i, pw, nbr, c, omega, vi, s, dt, strength, stride, out);
    }
    resolve_velocity(ld4(x, i), ld4(vsnap, i), out, colliders_of<kBatched>(col, i, st), radius, e,
                     f);
    st4(vout, i, out);
    });
}

// Clears words with a kernel rather than a memset node, so the step stays
// capturable in a graph.
__global__ void __launch_bounds__(kBlock) k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// CollisionSystem::solve (collision_system.cu:133-172) in one launch.
__global__ void __launch_bounds__(kBlock)
    k_collide_fixed_point(pbf_float4* __restrict__ p, size_t n, Colliders col, float radius,
                          uint32_t* __restrict__ status) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 pi = ld4(p, i);
    if (!collision_fixed_point(pi, col, radius))
        atomicOr(status, PBF_STATUS_COLLISION_NONCONVERGENCE);
    st4(p, i, pi);
}

// resolveVelocitiesKernel as the last pass when confinement is off.
template <bool kBatched>
__global__ void __launch_bounds__(kBlock)
    k_resolve_guarded(const pbf_float4* __restrict__ x, const pbf_float4* __restrict__ vsnap,
                      pbf_float4* __restrict__ v, size_t n, Colliders col, float r, float e,
                      float f, pbfint::SceneTable st, const uint32_t* __restrict__ status) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (*status != 0u) return;
    float4 vi = ld4(v, i);
    resolve_velocity(ld4(x, i), ld4(vsnap, i), vi, colliders_of<kBatched>(col, i, st), r, e, f);
    st4(v, i, vi);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI launchers (reference kernels)
// ---------------------------------------------------------------------------
extern "C" {

pbf_status pbf_k_integrate(const pbf_float4* x, pbf_float4* p, pbf_float4* v, int n, float dt,
                           pbf_float3 g, void* stream) {
    if (n <= 0) return PBF_OK;
    PBF_LAUNCH(k_integrate, grid_for(n), kBlock, stream, x, p, v, n, dt, g);
}

pbf_status pbf_k_compute_lambda(const pbf_float4* p, const uint32_t* nbr, const int* cnt,
                                int maxN, size_t n, float h, float mass, float rho0, float eps,
                                float* lambda, size_t stride, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_lambda_list<false>, grid_for(n), kBlock, stream, p, nbr, cnt, maxN, n, h, mass,
               rho0, eps, lambda, size_t{1}, stride, static_cast<uint32_t*>(nullptr), 0, 1,
               static_cast<pbf_float4*>(nullptr));
}

pbf_status pbf_k_compute_delta_position(const pbf_float4* p, const uint32_t* nbr, const int* cnt,
                                        int maxN, const float* lambda, size_t n, float h,
                                        float mass, float rho0, float sk, int sn, float sdq,
                                        pbf_float4* delta, size_t stride, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_delta_list, grid_for(n), kBlock, stream, p, nbr, cnt, maxN, lambda, n, h, mass,
               rho0, sk, sn, sdq, delta, stride);
}

pbf_status pbf_k_apply_delta_position(pbf_float4* p, const pbf_float4* d, size_t n, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_apply_delta, grid_for(n), kBlock, stream, p, d, n);
}

pbf_status pbf_k_update_velocity_and_position(pbf_float4* x, const pbf_float4* p, pbf_float4* v,
                                              size_t n, float inv_dt, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_update_velocity, grid_for(n), kBlock, stream, x, p, v, n, inv_dt);
}

pbf_status pbf_k_apply_xsph_viscosity(const pbf_float4* p, const uint32_t* nbr, const int* cnt,
                                      int maxN, const pbf_float4* vin, pbf_float4* vout, size_t n,
                                      float h, float visc, size_t stride, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_xsph_list, grid_for(n), kBlock, stream, p, nbr, cnt, maxN, vin, vout, n, h, visc,
               stride);
}

pbf_status pbf_k_compute_vorticity(const pbf_float4* p, const pbf_float4* v, const uint32_t* nbr,
                                   const int* cnt, int maxN, size_t n, float h, pbf_float4* omega,
                                   size_t stride, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_vorticity_list<false>, grid_for(n), kBlock, stream, p, v, nbr, cnt, maxN, n, h,
               omega, static_cast<pbf_float4*>(nullptr), stride, static_cast<uint32_t*>(nullptr),
               0, 1, static_cast<pbf_float4*>(nullptr));
}

pbf_status pbf_k_apply_vorticity_confinement(const pbf_float4* p, const uint32_t* nbr,
                                             const int* cnt, int maxN, const pbf_float4* omega,
                                             const pbf_float4* vin, pbf_float4* vout, size_t n,
                                             float h, float dt, float strength, size_t stride,
                                             void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_confine_list, grid_for(n), kBlock, stream, p, nbr, cnt, maxN, omega, vin, vout,
               n, h, dt, strength, stride);
}

pbf_status pbf_k_gather_solver_state(const uint32_t* idx, const pbf_float4* x, const pbf_float4* p,
                                     const pbf_float4* v, const pbf_float4* vin,
                                     const uint32_t* id, pbf_float4* xo, pbf_float4* po,
                                     pbf_float4* vo, pbf_float4* vino, uint32_t* ido, size_t n,
                                     void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_gather, grid_for(n), kBlock, stream, idx, x, p, v, vin, id, xo, po, vo, vino, ido,
               n);
}

pbf_status pbf_k_initialize_stable_particle_ids(uint32_t* id, size_t n, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_init_ids, grid_for(n), kBlock, stream, id, n);
}

pbf_status pbf_k_scatter_float4_by_stable_id(const pbf_float4* vals, const uint32_t* id,
                                             pbf_float4* out, size_t n, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_scatter_by_id, grid_for(n), kBlock, stream, vals, id, out, n);
}

pbf_status pbf_k_solve_container(pbf_float4* p, size_t n, const pbf_container* c, float r,
                                 int* flag, void* stream) {
    if (n == 0) return PBF_OK;
    if (p == nullptr || c == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null pointer");
    PBF_LAUNCH(k_solve_container, grid_for(n), kBlock, stream, p, n, c, r, flag);
}

pbf_status pbf_k_solve_spheres(pbf_float4* p, size_t n, const pbf_sphere* s, size_t count, float r,
                               int* flag, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_solve_spheres, grid_for(n), kBlock, stream, p, n, s, count, r, flag);
}

pbf_status pbf_k_solve_boxes(pbf_float4* p, size_t n, const pbf_box* b, size_t count, float r,
                             int* flag, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_solve_boxes, grid_for(n), kBlock, stream, p, n, b, count, r, flag);
}

pbf_status pbf_k_solve_planes(pbf_float4* p, size_t n, const pbf_plane* pl, size_t count, float r,
                              int* flag, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_solve_planes, grid_for(n), kBlock, stream, p, n, pl, count, r, flag);
}

pbf_status pbf_k_resolve_velocities(const pbf_float4* x, const pbf_float4* vin, pbf_float4* v,
                                    size_t n, const pbf_container* c, const pbf_sphere* s,
                                    size_t ns, const pbf_box* b, size_t nb, const pbf_plane* pl,
                                    size_t np, float r, float e, float f, void* stream) {
    if (n == 0) return PBF_OK;
    if (c == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "container must not be null");
    const Colliders col{c, s, b, pl, static_cast<int>(ns), static_cast<int>(nb),
                        static_cast<int>(np)};
    PBF_LAUNCH(k_resolve_velocities, grid_for(n), kBlock, stream, x, vin, v, n, col, r, e, f);
}

pbf_status pbf_k_poly6(const pbf_float3* d, size_t n, float h, float* out, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_poly6, grid_for(n), kBlock, stream, d, n, h, out);
}

pbf_status pbf_k_spiky_gradient(const pbf_float3* d, size_t n, float h, pbf_float3* out,
                                void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_spiky, grid_for(n), kBlock, stream, d, n, h, out);
}

pbf_status pbf_k_artificial_pressure(const float* d2, size_t n, float h, float k, int sn, float dq,
                                     float* out, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_scorr, grid_for(n), kBlock, stream, d2, n, h, k, sn, dq, out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Internal launchers of the fused step kernels (pbf_internal.hpp)
// ---------------------------------------------------------------------------
namespace pbfint {

pbf_status launch_predict_key(const pbf_float4* x, pbf_float4* p, pbf_float4* v, pbf_float4* vin,
                              uint32_t* key, int n, float dt, pbf_float3 g, pbf_float3 mn,
                              float cs, pbf_int3 gs, const pbf_container* box, float radius,
                              const SceneTable& st, int* clearCells, size_t clearCount,
                              const uint32_t* status, void* stream) {
    if (st.geom != nullptr)
        PBF_LAUNCH(k_predict_key<true>, grid_for(n), kBlock, stream, x, p, v, vin, key, n, dt, g,
                   mn, cs, gs, box, radius, st, clearCells, clearCount, status);
    PBF_LAUNCH(k_predict_key<false>, grid_for(n), kBlock, stream, x, p, v, vin, key, n, dt, g, mn,
               cs, gs, box, radius, st, clearCells, clearCount, status);
}

pbf_status launch_checkpoint(const pbf_float4* x, const pbf_float4* v, const uint32_t* id,
                             pbf_float4* cx, pbf_float4* cv, uint32_t* cid, size_t n,
                             const uint32_t* status, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_checkpoint, grid_for(n), kBlock, stream, x, v, id, cx, cv, cid, n, status);
}

pbf_status launch_frame_done(uint32_t* status, void* stream) {
    PBF_LAUNCH(k_frame_done, 1, 64, stream, status);
}

pbf_status launch_collide_fixed_point(pbf_float4* p, size_t n, const Colliders& col, float radius,
                                      uint32_t* status, void* stream) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_collide_fixed_point, grid_for(n), kBlock, stream, p, n, col, radius, status);
}

pbf_status launch_zero_u32(uint32_t* p, size_t n, void* stream) {
    PBF_LAUNCH(k_zero_u32, grid_for(n), kBlock, stream, p, n);
}

pbf_status launch_keys(const pbf_float4* p, uint32_t* key, size_t n, pbf_float3 mn, float cs,
                       pbf_int3 gs, const SceneTable& st, void* stream) {
    if (st.geom != nullptr)
        PBF_LAUNCH(k_keys<true>, grid_for(n), kBlock, stream, p, key, n, mn, cs, gs, st);
    PBF_LAUNCH(k_keys<false>, grid_for(n), kBlock, stream, p, key, n, mn, cs, gs, st);
}

pbf_status launch_lambda_guarded(pbf_float4* p, const uint32_t* nbr, const int* cnt, int maxN,
                                 size_t n, float h, float mass, float rho0, float eps,
                                 size_t stride, const ListPass& lp, uint32_t* status, void* stream) {
    float* lambda = reinterpret_cast<float*>(p) + 3;  // the .w words
    if (lp.p0 != nullptr)
        PBF_LAUNCH((k_lambda_list<true, true>), grid_for(n, lp.block * lp.ppt), lp.block, stream, p, nbr, cnt, maxN, n,
                   h, mass, rho0, eps, lambda, size_t{4}, stride, status, lp.exactMax, lp.ppt,
                   lp.tapeOut);
    PBF_LAUNCH((k_lambda_list<true, false>), grid_for(n, lp.block * lp.ppt), lp.block, stream, p, nbr, cnt, maxN, n, h,
               mass, rho0, eps, lambda, size_t{4}, stride, status, 0, lp.ppt, lp.tapeOut);
}

pbf_status launch_delta_apply_collide(const pbf_float4* pin, pbf_float4* pout, const uint32_t* nbr,
                                      const int* cnt, int maxN, const pbf_float4* x, size_t n,
                                      float h, float mass, float rho0, float sk, int sn, float sdq,
                                      size_t stride, const Colliders& col, float radius,
                                      const SceneTable& st, const ListPass& lp, uint32_t* status,
                                      void* stream) {
    const bool batched = st.geom != nullptr;
    const bool skinned = lp.p0 != nullptr;
#define PBF_DELTA(B, S)                                                                     \
    PBF_LAUNCH((k_delta_apply_collide<B, S>), grid_for(n, lp.block * lp.ppt), lp.block, stream, pin,  \
               pout, nbr, cnt, maxN, x, n, h, mass, rho0, sk, sn, sdq, stride, col, radius, \
               st, lp.p0, lp.stale2, lp.keyOut, lp.keyMin, lp.keyCell, lp.keyDims,          \
               lp.clearCells, lp.clearCount, lp.xperm, lp.prePos, status, lp.ppt)
    if (batched && skinned) PBF_DELTA(true, true);
    if (batched) PBF_DELTA(true, false);
    if (skinned) PBF_DELTA(false, true);
    PBF_DELTA(false, false);
#undef PBF_DELTA
}

pbf_status launch_update_xsph(const pbf_float4* x, const pbf_float4* p, const pbf_float4* vold,
                              pbf_float4* vout, const uint32_t* nbr, const int* cnt, int maxN,
                              size_t n, float h, float visc, float inv_dt, size_t stride,
                              const ListPass& lp, uint32_t* status, void* stream) {
    if (lp.exactMax > 0)
        PBF_LAUNCH(k_update_xsph<true>, grid_for(n, lp.block * lp.ppt), lp.block, stream, x, p, vold, vout, nbr, cnt,
                   maxN, n, h, visc, inv_dt, stride, status, lp.exactMax, lp.ppt, lp.tapeOut);
    PBF_LAUNCH(k_update_xsph<false>, grid_for(n, lp.block * lp.ppt), lp.block, stream, x, p, vold, vout, nbr, cnt, maxN,
               n, h, visc, inv_dt, stride, status, 0, lp.ppt, lp.tapeOut);
}

pbf_status launch_update_velocity_only(const pbf_float4* x, const pbf_float4* p, pbf_float4* v,
                                       size_t n, float inv_dt, const uint32_t* status,
                                       void* stream) {
    PBF_LAUNCH(k_update_velocity_only, grid_for(n), kBlock, stream, x, p, v, n, inv_dt, status);
}

pbf_status launch_vorticity(const pbf_float4* p, const pbf_float4* v, const uint32_t* nbr,
                            const int* cnt, int maxN, size_t n, float h, pbf_float4* omega,
                            pbf_float4* pw, size_t stride, const ListPass& lp, uint32_t* status,
                            void* stream) {
    PBF_LAUNCH(k_vorticity_list<true>, grid_for(n, lp.block * lp.ppt), lp.block, stream, p, v, nbr, cnt, maxN, n, h,
               omega, pw, stride, status, lp.exactMax, lp.ppt, lp.tapeOut);
}

pbf_status launch_confine_resolve(const pbf_float4* x, const pbf_float4* pw, const uint32_t* nbr,
                                  const int* cnt, int maxN, const pbf_float4* omega,
                                  const pbf_float4* vcur, const pbf_float4* vsnap, pbf_float4* vout,
                                  size_t n, float h, float dt, float strength, size_t stride,
                                  const Colliders& col, float radius, float e, float f,
                                  const SceneTable& st, const ListPass& lp,
                                  const uint32_t* status, void* stream) {
    if (st.geom != nullptr)
        PBF_LAUNCH(k_confine_resolve<true>, grid_for(n, lp.block * lp.ppt), lp.block, stream, x, pw, nbr, cnt, maxN,
                   omega, vcur, vsnap, vout, n, h, dt, strength, stride, col, radius, e, f, st,
                   status, lp.ppt);
    PBF_LAUNCH(k_confine_resolve<false>, grid_for(n, lp.block * lp.ppt), lp.block, stream, x, pw, nbr, cnt, maxN, omega,
               vcur, vsnap, vout, n, h, dt, strength, stride, col, radius, e, f, st, status,
               lp.ppt);
}

pbf_status launch_resolve_guarded(const pbf_float4* x, const pbf_float4* vsnap, pbf_float4* v,
                                  size_t n, const Colliders& col, float r, float e, float f,
                                  const SceneTable& st, const uint32_t* status, void* stream) {
    if (st.geom != nullptr)
        PBF_LAUNCH(k_resolve_guarded<true>, grid_for(n), kBlock, stream, x, vsnap, v, n, col, r, e,
                   f, st, status);
    PBF_LAUNCH(k_resolve_guarded<false>, grid_for(n), kBlock, stream, x, vsnap, v, n, col, r, e, f,
               st, status);
}

}  // namespace pbfint
