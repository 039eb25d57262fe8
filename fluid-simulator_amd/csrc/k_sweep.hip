// LDS-tiled neighbour sweeps of the solver step (gfx950).
//
// The reference writes a neighbour list per particle (src/pbf/neighbors.cu)
// and every constraint and velocity kernel gathers through it
// (src/pbf/pbf_solver_kernels.cu).  Here each neighbour pass sweeps the
// cells itself, with the candidates staged in LDS:
//
//   * a tile is up to 32 consecutive cells of one grid row (fixed y, z).
//     Particles are sorted by cell key (x fastest), so the cells
//     x0-1 .. x1+1 of any row are one contiguous range of the working set:
//     a tile's halo is 9 such row segments (dz, dy in -1..1).  Sorted keys
//     visit tiles in increasing order; workgroup b takes the tiles whose
//     first particle lies in [256 b, 256 b + 256), so no empty tile is ever
//     visited and no list of tiles is built;
//   * the 9 segments are staged one z-plane (3 rows) at a time with
//     coalesced loads, in the reference's z, y, x stencil order;
//   * each thread owns one particle of the tile.  For every row it tests the
//     candidates of cells cx-1 .. cx+1 against the inclusive FP32 radius
//     d2 <= h*h (no FMA) into 32-bit masks, then evaluates the pass's
//     physics for the set bits in increasing order.
//
// The neighbour set and order of every particle are exactly those of the
// reference list (same candidates, same radius test, same order); the
// count is checked against the list capacity for the overflow error.
//
// The first pass after a grid build (the producer) stores each particle's
// masks, two 32-bit words per row; the passes that follow on the same grid
// and positions (the consumers) read them instead of testing again.  A row
// with more than 64 candidates is not stored and is tested by every pass.
// A plane whose staged halo exceeds the LDS capacity is read from global
// memory instead, in the same order.  A workgroup whose range starts more
// than kMaxTilePasses tiles (sparse regions: spray, surface) gives each of
// their particles a lane of its own that reads its 27 cells from global
// memory, rather than sweeping the tiles one after another; the candidate
// rows and mask bits are the same in both forms.
//
// Requires a 3x3x3 stencil (cell size == smoothing radius), which the
// solver's grid guarantees (src/pbf/pbf_solver.cu:81-82).
#include "pbf_device.hpp"
#include "pbf_internal.hpp"
#include "pbf_launch.hpp"

using namespace pbfdev;

namespace {

constexpr int kTileW = static_cast<int>(pbfint::kSweepTileW);  // cells per tile along x
constexpr int kSegCells = kTileW + 2;  // tile cells plus the x halo
constexpr int kSegs = 9;               // (dz, dy) rows, dz-major
constexpr int kThreads = 256;
constexpr int kCap = 1536;             // staged particles per z-plane
constexpr int kPad = 4;                // test reads run up to 3 records past a row
constexpr int kRowWords = 2;           // stored mask words per row (64 candidates)
constexpr int kMaxTilePasses = 4;      // more tiles per workgroup take the lane path

struct Grid {
    const int* cellStart;
    const int* cellEnd;
    const uint32_t* sortedKeys;
    const uint32_t* tileEnd;
    uint32_t* masks;  // [kSegs * kRowWords][stride]
    size_t n;
    size_t stride;
    pbf_int3 dims;
    int chunks;  // tiles per row = ceil(dims.x / kTileW)
    float h;
    float h2;
    int maxNeighbors;
};

struct Meta {
    int segFirst[kSegs];  // working-set index of the segment's first particle
    int segBase[kSegs];   // LDS slot of that particle within its plane
    int segCells[kSegs];  // cells of the segment (0: empty or outside the grid)
    int planeCount[3];
    int cellOff[kSegs][kSegCells + 1];  // exclusive prefix of per-cell counts
    int stop;
    int waveTiles[kThreads / 64];
    int tiles[kThreads];       // tiles starting in this workgroup's particle range
    int tileFirst[kThreads];   // their first particles
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8).  Give every XCD one contiguous run of particle ranges so
// neighbouring tiles, which share halo rows, share that XCD's L2.
__device__ __forceinline__ int xcd_task(int b, int nb) {
    const int xcd = b & 7;
    const int k = b >> 3;
    const int per = nb >> 3;
    const int rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// Workgroup barrier for LDS hand-offs only.  No thread of a sweep reads
// global memory that another thread of the same launch writes, so the
// barrier orders LDS alone and leaves global loads in flight across it
// (__syncthreads would wait for every outstanding global load first).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t tile_of(uint32_t key, uint32_t gx, uint32_t chunks) {
    return (key / gx) * chunks + (key % gx) / static_cast<uint32_t>(kTileW);
}

// Staged records of one z-plane, one array per component: lanes of a wave
// sit in ~8 neighbouring cells and read records ~8 apart at every step,
// which 16-byte records would map onto one bank group; 4-byte components
// spread them over the banks, and the three position loads of a candidate
// share one address register with immediate offsets.
template <bool kExt>
struct Plane {
    float x[kCap + kPad];
    float y[kCap + kPad];
    float z[kCap + kPad];
    float w[kCap + kPad];
    float e0[kExt ? kCap : 1];
    float e1[kExt ? kCap : 1];
};

// Candidate records of one row segment: staged in LDS, or read from the
// working set when the plane did not fit.  Index q is relative to the
// segment's first particle; position() may be asked for up to 3 records
// past the row (padding in LDS, clamped in global memory).
template <class Pass>
struct LdsRow {
    const Plane<Pass::kExt>* pl;
    int base;
    __device__ float4 position(int q) const {
        const int k = base + q;
        return make_float4(pl->x[k], pl->y[k], pl->z[k], 0.0f);
    }
    __device__ void record(int q, float4& p, float2& e) const {
        const int k = base + q;
        p = make_float4(pl->x[k], pl->y[k], pl->z[k], pl->w[k]);
        if (Pass::kExt) e = make_float2(pl->e0[k], pl->e1[k]);
    }
};

template <class Pass>
struct GlobalRow {
    const typename Pass::Args* a;
    size_t first;
    int last;
    __device__ float4 position(int q) const {
        float4 p;
        float2 e;
        Pass::stage(*a, first + static_cast<size_t>(min(q, last)), p, e);
        return p;
    }
    __device__ void record(int q, float4& p, float2& e) const {
        Pass::stage(*a, first + static_cast<size_t>(q), p, e);
    }
};

// Neighbours of one particle among candidates [lo, hi) of one row, in
// order, in windows of 32 candidates.  A window's mask comes from the
// radius test (branch-free, four loads in flight) or, for a consumer, from
// the producer's stored words.  The physics then runs over the set bits,
// lowest first, with the next record loaded ahead.
template <class Pass, bool kConsume, class Src>
__device__ __forceinline__ void sweep_row(const typename Pass::Args& pa, const Src& src, float4 ip,
                                          float2 ie, int lo, int hi, int selfq, float h2,
                                          typename Pass::Acc& acc, int& count, uint32_t* words,
                                          size_t stride, uint32_t pre0, uint32_t pre1) {
    const bool kept = hi - lo <= 32 * kRowWords;
    int k = 0;
    for (int c = lo; c < hi; c += 32, ++k) {
        uint32_t mask;
        if (kConsume && kept) {
            mask = k == 0 ? pre0 : pre1;
        } else {
            const int w = min(32, hi - c);
            mask = 0u;
            for (int b = 0; b < w; b += 4) {
                float4 j[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) j[u] = src.position(c + b + u);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float d2 = dist2_exact(ip.x - j[u].x, ip.y - j[u].y, ip.z - j[u].z);
                    mask |= static_cast<uint32_t>(d2 <= h2) << (b + u);
                }
            }
            if (w < 32) mask &= (1u << w) - 1u;
            if (selfq >= c && selfq < c + 32) mask &= ~(1u << (selfq - c));
            if (!kConsume) {
                count += __popc(mask);
                if (kept) words[static_cast<size_t>(k) * stride] = mask;
            }
        }
        if (mask == 0u) continue;
        int b = __builtin_ctz(mask);
        mask &= mask - 1u;
        float4 jp;
        float2 je = make_float2(0.0f, 0.0f);
        src.record(c + b, jp, je);
        for (;;) {
            const bool more = mask != 0u;
            const int nb = more ? __builtin_ctz(mask) : b;
            float4 np;
            float2 ne = make_float2(0.0f, 0.0f);
            src.record(c + nb, np, ne);
            const float dx = ip.x - jp.x, dy = ip.y - jp.y, dz = ip.z - jp.z;
            Pass::add(pa, acc, ip, ie, jp, je, dx, dy, dz, dist2_exact(dx, dy, dz));
            if (!more) break;
            mask &= mask - 1u;
            b = nb;
            jp = np;
            je = ne;
        }
    }
}

// One tile: segment geometry, then the tile's particles against the three
// z-planes of its halo.  Every branch that reaches a barrier is uniform
// over the workgroup (it depends on LDS values written before a barrier).
template <class Pass, bool kConsume>
__device__ __forceinline__ void sweep_tile(const Grid& g, const typename Pass::Args& pa, Meta& m,
                                           Plane<Pass::kExt>& lds, int tile) {
    const int row = tile / g.chunks;
    const int chunk = tile - row * g.chunks;
    const int y = row % g.dims.y;
    const int z = row / g.dims.y;
    const int x0 = chunk * kTileW;
    const int x1 = min(x0 + kTileW, g.dims.x) - 1;
    const int xs = max(x0 - 1, 0);
    const int xe = min(x1 + 1, g.dims.x - 1);
    const int ncell = xe - xs + 1;
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // Segment geometry: one wave per row, one lane per cell.
    for (int s = wave; s < kSegs; s += kThreads / 64) {
        const int yy = y + (s % 3) - 1;
        const int zz = z + (s / 3) - 1;
        const bool rowOk = yy >= 0 && yy < g.dims.y && zz >= 0 && zz < g.dims.z;
        int cnt = 0;
        int st = 0x7fffffff;
        if (rowOk && lane < ncell) {
            const uint32_t key = key_of(xs + lane, yy, zz, g.dims);
            const int a = g.cellStart[key];
            if (a != -1) {
                cnt = g.cellEnd[key] - a;
                st = a;
            }
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        int first = st;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) first = min(first, __shfl_xor(first, o, 64));
        if (lane <= ncell) m.cellOff[s][lane] = incl - cnt;
        if (lane == 63) {
            m.segFirst[s] = incl > 0 ? first : 0;
            m.segCells[s] = incl > 0 ? ncell : 0;
        }
    }
    lds_barrier();
    if (tid < 3) {
        int base = 0;
        for (int r = 0; r < 3; ++r) {
            const int s = 3 * tid + r;
            m.segBase[s] = base;
            base += m.segCells[s] > 0 ? m.cellOff[s][m.segCells[s]] : 0;
        }
        m.planeCount[tid] = base;
    }
    lds_barrier();

    const int c0 = m.cellOff[4][x0 - xs];
    const int c1 = m.cellOff[4][x1 - xs + 1];
    const size_t own0 = static_cast<size_t>(m.segFirst[4]);

    // Usually one round; more when the tile holds more particles than threads.
    for (int r0 = c0; r0 < c1; r0 += kThreads) {
        const int r = r0 + tid;
        const bool active = r < c1;
        const size_t gi = own0 + static_cast<size_t>(active ? r : c0);
        float4 ip = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float2 ie = make_float2(0.0f, 0.0f);
        int mc = 0;
        typename Pass::Acc acc{};
        int count = 0;
        if (active) {
            Pass::stage(pa, gi, ip, ie);
            mc = static_cast<int>(g.sortedKeys[gi] % static_cast<uint32_t>(g.dims.x)) - xs;
            acc = Pass::begin(pa, ip, ie);
        }
        for (int pl = 0; pl < 3; ++pl) {
            const int pc = m.planeCount[pl];
            if (pc == 0) continue;
            const bool staged = pc <= kCap;
            // Row ranges of this plane, and for a consumer its stored mask
            // words, requested before the staging so their latency overlaps.
            int lo[3], hi[3];
            uint32_t pre[3][2];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int s = 3 * pl + rr;
                const int nc = m.segCells[s];
                lo[rr] = 0;
                hi[rr] = 0;
                pre[rr][0] = 0u;
                pre[rr][1] = 0u;
                if (!active || nc == 0) continue;
                lo[rr] = m.cellOff[s][max(mc - 1, 0)];
                hi[rr] = m.cellOff[s][min(mc + 1, nc - 1) + 1];
                const int len = hi[rr] - lo[rr];
                if (kConsume && len > 0 && len <= 32 * kRowWords) {
                    const uint32_t* w = g.masks + static_cast<size_t>(s * kRowWords) * g.stride + gi;
                    pre[rr][0] = w[0];
                    if (len > 32) pre[rr][1] = w[g.stride];
                }
            }
            if (staged) {
                lds_barrier();  // the previous plane's readers are done
                for (int rr = 0; rr < 3; ++rr) {
                    const int s = 3 * pl + rr;
                    if (m.segCells[s] == 0) continue;
                    const int cnt = m.cellOff[s][m.segCells[s]];
                    const size_t first = static_cast<size_t>(m.segFirst[s]);
                    const int base = m.segBase[s];
                    for (int q = tid; q < cnt; q += kThreads) {
                        float4 a;
                        float2 b;
                        Pass::stage(pa, first + q, a, b);
                        const int k = base + q;
                        lds.x[k] = a.x;
                        lds.y[k] = a.y;
                        lds.z[k] = a.z;
                        lds.w[k] = a.w;
                        if (Pass::kExt) {
                            lds.e0[k] = b.x;
                            lds.e1[k] = b.y;
                        }
                    }
                }
                lds_barrier();
            }
            if (!active) continue;
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int s = 3 * pl + rr;
                if (hi[rr] <= lo[rr]) continue;
                const int selfq = s == 4 ? r : -1;
                uint32_t* words = g.masks + static_cast<size_t>(s * kRowWords) * g.stride + gi;
                if (staged) {
                    const LdsRow<Pass> src{&lds, m.segBase[s]};
                    sweep_row<Pass, kConsume>(pa, src, ip, ie, lo[rr], hi[rr], selfq, g.h2, acc,
                                              count, words, g.stride, pre[rr][0], pre[rr][1]);
                } else {
                    const GlobalRow<Pass> src{&pa, static_cast<size_t>(m.segFirst[s]),
                                              m.cellOff[s][m.segCells[s]] - 1};
                    sweep_row<Pass, kConsume>(pa, src, ip, ie, lo[rr], hi[rr], selfq, g.h2, acc,
                                              count, words, g.stride, pre[rr][0], pre[rr][1]);
                }
            }
        }
        if (active) {
            if (!kConsume && count > g.maxNeighbors)
                atomicOr(pa.status, PBF_STATUS_NEIGHBOR_OVERFLOW);
            Pass::end(pa, acc, gi, ip, ie);
        }
    }
    lds_barrier();  // LDS is reused by the next tile
}

// One particle on one lane, candidates read from global memory: the row
// range of every (dz, dy) is the run of particles from the first non-empty
// cell of cx-1 .. cx+1 to the last, as in the tile form.
template <class Pass, bool kConsume>
__device__ __forceinline__ void sweep_lane(const Grid& g, const typename Pass::Args& pa,
                                           size_t gi) {
    const uint32_t key = g.sortedKeys[gi];
    const uint32_t gx = static_cast<uint32_t>(g.dims.x);
    const int cx = static_cast<int>(key % gx);
    const uint32_t rowId = key / gx;
    const int cy = static_cast<int>(rowId % static_cast<uint32_t>(g.dims.y));
    const int cz = static_cast<int>(rowId / static_cast<uint32_t>(g.dims.y));
    float4 ip;
    float2 ie = make_float2(0.0f, 0.0f);
    Pass::stage(pa, gi, ip, ie);
    typename Pass::Acc acc = Pass::begin(pa, ip, ie);
    int count = 0;
    const int xa = max(cx - 1, 0);
    const int xb = min(cx + 1, g.dims.x - 1);
    for (int s = 0; s < kSegs; ++s) {
        const int yy = cy + (s % 3) - 1;
        const int zz = cz + (s / 3) - 1;
        if (yy < 0 || yy >= g.dims.y || zz < 0 || zz >= g.dims.z) continue;
        int rowStart = -1, rowEnd = -1;
        for (int x = xa; x <= xb; ++x) {
            const uint32_t k = key_of(x, yy, zz, g.dims);
            const int a = g.cellStart[k];
            if (a == -1) continue;
            if (rowStart < 0) rowStart = a;
            rowEnd = g.cellEnd[k];
        }
        if (rowStart < 0) continue;
        const int len = rowEnd - rowStart;
        const int selfq = s == 4 ? static_cast<int>(gi - static_cast<size_t>(rowStart)) : -1;
        const GlobalRow<Pass> src{&pa, static_cast<size_t>(rowStart), len - 1};
        uint32_t* words = g.masks + static_cast<size_t>(s * kRowWords) * g.stride + gi;
        uint32_t pre0 = 0u, pre1 = 0u;
        if (kConsume && len <= 32 * kRowWords) {
            pre0 = words[0];
            if (len > 32) pre1 = words[g.stride];
        }
        sweep_row<Pass, kConsume>(pa, src, ip, ie, 0, len, selfq, g.h2, acc, count, words,
                                  g.stride, pre0, pre1);
    }
    if (!kConsume && count > g.maxNeighbors) atomicOr(pa.status, PBF_STATUS_NEIGHBOR_OVERFLOW);
    Pass::end(pa, acc, gi, ip, ie);
}

// One workgroup per 256 sorted particles: it sweeps every tile whose first
// particle falls in its range (none, one, or several small ones).
template <class Pass, bool kConsume>
__global__ void __launch_bounds__(kThreads) k_sweep(Grid g, typename Pass::Args args) {
    __shared__ Meta m;
    __shared__ Plane<Pass::kExt> lds;

    // SPH constants are evaluated on the device, as in the other kernels.
    typename Pass::Args pa = args;
    pa.s = sph_consts(g.h);
    Pass::prepare(pa);

    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const size_t i = static_cast<size_t>(xcd_task(static_cast<int>(blockIdx.x),
                                                  static_cast<int>(gridDim.x))) *
                         kThreads + static_cast<size_t>(tid);
    const uint32_t gx = static_cast<uint32_t>(g.dims.x);
    const uint32_t chunks = static_cast<uint32_t>(g.chunks);
    bool starts = false;
    uint32_t t = 0;
    if (i < g.n) {
        t = tile_of(g.sortedKeys[i], gx, chunks);
        starts = i == 0 || tile_of(g.sortedKeys[i - 1], gx, chunks) != t;
    }
    // Tiles in particle order: wave counts, then ranks within the wave.
    const uint64_t ballot = __ballot(starts);
    if (lane == 0) m.waveTiles[wave] = __popcll(ballot);
    // A raised status word stops the step; read once, through LDS, so the
    // whole workgroup takes the decision together.
    if (tid == 0) m.stop = *pa.status != 0u ? 1 : 0;
    lds_barrier();
    if (m.stop != 0) return;
    int before = 0, total = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        before += w < wave ? m.waveTiles[w] : 0;
        total += m.waveTiles[w];
    }
    if (starts) {
        const int rank = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
            static_cast<uint32_t>(ballot >> 32),
            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ballot), 0u)));
        m.tiles[before + rank] = static_cast<int>(t);
        m.tileFirst[before + rank] = static_cast<int>(i);
    }
    lds_barrier();
    if (total <= kMaxTilePasses) {
        for (int k = 0; k < total; ++k)
            sweep_tile<Pass, kConsume>(g, pa, m, lds, m.tiles[k]);
    } else {
        // The particles of every tile that starts here, one lane each.
        const size_t first = static_cast<size_t>(m.tileFirst[0]);
        const size_t end = g.tileEnd[m.tiles[total - 1]];
        for (size_t q = first + static_cast<size_t>(tid); q < end; q += kThreads)
            sweep_lane<Pass, kConsume>(g, pa, q);
    }
}

// ---------------------------------------------------------------------------
// Passes.  stage() yields the LDS record of working-set particle g: a float4
// whose xyz is the position, plus a float2 when kExt.  begin/add/end are the
// per-particle accumulation of the reference kernel the pass replaces.
// ---------------------------------------------------------------------------

// computeLambda (src/pbf/pbf_solver_kernels.cu:26-105): per-neighbour terms
// in FP32, their sums in FP64 (C = rho / rho0 - 1 cancels most of an FP32
// sum's significant bits near the rest density).
struct LambdaPass {
    static constexpr bool kExt = false;
    struct Args {
        const pbf_float4* p;
        float* lambda;
        uint32_t* status;
        Sph s;
        float mass, rho0, eps, gs;
    };
    struct Acc {
        double rho, gx, gy, gz, sg;
    };
    __device__ static void prepare(Args&) {}
    __device__ static void stage(const Args& a, size_t g, float4& pos, float2&) {
        pos = ld4(a.p, g);
    }
    __device__ static Acc begin(const Args& a, float4, float2) {
        return Acc{static_cast<double>(a.mass * a.s.poly_norm), 0.0, 0.0, 0.0, 0.0};
    }
    __device__ static void add(const Args& a, Acc& c, float4, float2, float4, float2, float dx,
                               float dy, float dz, float d2) {
        c.rho += static_cast<double>(a.mass * poly6_d2(d2, a.s));
        const float3 gw = spiky_grad_rsq(dx, dy, dz, d2, a.s);
        const float gjx = -a.gs * gw.x, gjy = -a.gs * gw.y, gjz = -a.gs * gw.z;
        c.sg += static_cast<double>(gjx * gjx + gjy * gjy + gjz * gjz);
        c.gx -= static_cast<double>(gjx);
        c.gy -= static_cast<double>(gjy);
        c.gz -= static_cast<double>(gjz);
    }
    __device__ static void end(const Args& a, Acc& c, size_t gi, float4, float2) {
        const double sg = c.sg + c.gx * c.gx + c.gy * c.gy + c.gz * c.gz;
        const double C = c.rho / static_cast<double>(a.rho0) - 1.0;
        a.lambda[gi] = static_cast<float>(-C / (sg + static_cast<double>(a.eps)));
    }
};

// computeDeltaPosition + applyDeltaPosition + CollisionSystem::solve
// (pbf_solver.cu:240-257).  The record carries lambda_j in .w; the output
// goes to a second buffer (Jacobi), so every particle sees the positions
// the reference's separate delta-p kernel sees.
struct DeltaPass {
    static constexpr bool kExt = false;
    struct Args {
        const pbf_float4* p;
        const float* lambda;
        pbf_float4* pout;
        uint32_t* status;
        Sph s;
        float sk;
        int sn;
        float sdq;
        Scorr sc;
        float scale;  // mass / rho0
        Colliders col;
        float radius;
    };
    struct Acc {
        float ax, ay, az;
    };
    __device__ static void prepare(Args& a) { a.sc = scorr_consts(a.s.h, a.sk, a.sn, a.sdq); }
    __device__ static void stage(const Args& a, size_t g, float4& pos, float2&) {
        pos = ld4(a.p, g);
        pos.w = a.lambda[g];
    }
    __device__ static Acc begin(const Args&, float4, float2) { return Acc{0.0f, 0.0f, 0.0f}; }
    __device__ static void add(const Args& a, Acc& c, float4 i, float2, float4 j, float2, float dx,
                               float dy, float dz, float d2) {
        const float3 gw = spiky_grad_rsq(dx, dy, dz, d2, a.s);
        const float ls = i.w + j.w + scorr_value(d2, a.s.h2, a.sc);
        c.ax += ls * gw.x;
        c.ay += ls * gw.y;
        c.az += ls * gw.z;
    }
    __device__ static void end(const Args& a, Acc& c, size_t gi, float4, float2) {
        float4 p = ld4(a.p, gi);
        p.x += a.scale * c.ax;
        p.y += a.scale * c.ay;
        p.z += a.scale * c.az;
        if (!collision_fixed_point(p, a.col, a.radius))
            atomicOr(a.status, PBF_STATUS_COLLISION_NONCONVERGENCE);
        st4(a.pout, gi, p);
    }
};

// updateVelocityAndPosition + applyXsphViscosity (pbf_solver.cu:284-298).
// v = (p* - x) / dt is rebuilt in the record (v.x in .w, v.y v.z in the
// float2) with the FP32 expression the reference stores; x <- p* is a
// buffer swap on the host.
struct XsphPass {
    static constexpr bool kExt = true;
    struct Args {
        const pbf_float4* x;
        const pbf_float4* p;
        const pbf_float4* vold;
        pbf_float4* vout;
        uint32_t* status;
        Sph s;
        float visc, invDt;
    };
    struct Acc {
        float cx, cy, cz;
    };
    __device__ static void prepare(Args&) {}
    __device__ static void stage(const Args& a, size_t g, float4& pos, float2& ext) {
        const float4 xg = ld4(a.x, g);
        const float4 pg = ld4(a.p, g);
        pos = make_float4(pg.x, pg.y, pg.z, (pg.x - xg.x) * a.invDt);
        ext = make_float2((pg.y - xg.y) * a.invDt, (pg.z - xg.z) * a.invDt);
    }
    __device__ static Acc begin(const Args&, float4, float2) { return Acc{0.0f, 0.0f, 0.0f}; }
    __device__ static void add(const Args& a, Acc& c, float4 i, float2 ie, float4 j, float2 je,
                               float, float, float, float d2) {
        const float w = poly6_d2(d2, a.s);
        c.cx += (j.w - i.w) * w;
        c.cy += (je.x - ie.x) * w;
        c.cz += (je.y - ie.y) * w;
    }
    __device__ static void end(const Args& a, Acc& c, size_t gi, float4 i, float2 ie) {
        const float vw = a.vold[gi].w;
        st4(a.vout, gi, make_float4(i.w + a.visc * c.cx, ie.x + a.visc * c.cy,
                                    ie.y + a.visc * c.cz, vw));
    }
};

// computeVorticity (src/pbf/pbf_solver_kernels.cu:268-315).
struct VorticityPass {
    static constexpr bool kExt = true;
    struct Args {
        const pbf_float4* p;
        const pbf_float4* v;
        pbf_float4* omega;
        uint32_t* status;
        Sph s;
    };
    struct Acc {
        float ox, oy, oz;
    };
    __device__ static void prepare(Args&) {}
    __device__ static void stage(const Args& a, size_t g, float4& pos, float2& ext) {
        const float4 pg = ld4(a.p, g);
        const float4 vg = ld4(a.v, g);
        pos = make_float4(pg.x, pg.y, pg.z, vg.x);
        ext = make_float2(vg.y, vg.z);
    }
    __device__ static Acc begin(const Args&, float4, float2) { return Acc{0.0f, 0.0f, 0.0f}; }
    __device__ static void add(const Args& a, Acc& c, float4 i, float2 ie, float4 j, float2 je,
                               float dx, float dy, float dz, float d2) {
        const float3 gr = spiky_grad_rsq(dx, dy, dz, d2, a.s);
        const float ux = j.w - i.w, uy = je.x - ie.x, uz = je.y - ie.y;
        c.ox += gr.y * uz - gr.z * uy;
        c.oy += gr.z * ux - gr.x * uz;
        c.oz += gr.x * uy - gr.y * ux;
    }
    __device__ static void end(const Args& a, Acc& c, size_t gi, float4, float2) {
        st4(a.omega, gi, make_float4(c.ox, c.oy, c.oz, 0.0f));
    }
};

// applyVorticityConfinement + resolveVelocitiesKernel (pbf_solver.cu:309-323).
// The record carries |omega_j| in .w.
struct ConfinePass {
    static constexpr bool kExt = false;
    struct Args {
        const pbf_float4* x;
        const pbf_float4* omega;
        const pbf_float4* vcur;
        const pbf_float4* vsnap;
        pbf_float4* vout;
        uint32_t* status;
        Sph s;
        float dt, strength;
        Colliders col;
        float radius, restitution, friction;
    };
    struct Acc {
        float ex, ey, ez;
    };
    __device__ static void prepare(Args&) {}
    __device__ static void stage(const Args& a, size_t g, float4& pos, float2&) {
        pos = ld4(a.x, g);
        const float4 w = ld4(a.omega, g);
        pos.w = __builtin_sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    }
    __device__ static Acc begin(const Args&, float4, float2) { return Acc{0.0f, 0.0f, 0.0f}; }
    __device__ static void add(const Args& a, Acc& c, float4 i, float2, float4 j, float2, float dx,
                               float dy, float dz, float d2) {
        const float3 gr = spiky_grad_rsq(dx, dy, dz, d2, a.s);
        const float md = j.w - i.w;
        c.ex += md * gr.x;
        c.ey += md * gr.y;
        c.ez += md * gr.z;
    }
    __device__ static void end(const Args& a, Acc& c, size_t gi, float4, float2) {
        const float4 vi = ld4(a.vcur, gi);
        const float4 wi = ld4(a.omega, gi);
        const float el = __builtin_sqrtf(c.ex * c.ex + c.ey * c.ey + c.ez * c.ez);
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if (el > 1.0e-6f) {  // normalizationEpsilon, pbf_solver_kernels.cu:367
            nx = c.ex / el;
            ny = c.ey / el;
            nz = c.ez / el;
        }
        const float fx = a.strength * (ny * wi.z - nz * wi.y);
        const float fy = a.strength * (nz * wi.x - nx * wi.z);
        const float fz = a.strength * (nx * wi.y - ny * wi.x);
        float4 out = make_float4(vi.x + a.dt * fx, vi.y + a.dt * fy, vi.z + a.dt * fz, vi.w);
        resolve_velocity(ld4(a.x, gi), ld4(a.vsnap, gi), out, a.col, a.radius, a.restitution,
                         a.friction);
        st4(a.vout, gi, out);
    }
};

template <class Pass, bool kConsume>
pbf_status launch_sweep(const pbfint::SweepGrid& sg, const typename Pass::Args& pa, void* stream) {
    if (sg.n == 0) return PBF_OK;
    Grid g;
    g.cellStart = sg.cellStart;
    g.cellEnd = sg.cellEnd;
    g.sortedKeys = sg.sortedKeys;
    g.tileEnd = sg.tileEnd;
    g.masks = sg.masks;
    g.n = sg.n;
    g.stride = sg.maskStride;
    g.dims = sg.dims;
    g.chunks = (sg.dims.x + kTileW - 1) / kTileW;
    g.h = sg.h;
    g.h2 = sg.h * sg.h;
    g.maxNeighbors = sg.maxNeighbors;
    const size_t tiles = static_cast<size_t>(g.chunks) * static_cast<size_t>(sg.dims.y) *
                         static_cast<size_t>(sg.dims.z);
    if (tiles > 0x7fffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "grid has too many rows for the neighbour sweep");
    if (sg.maskStride < sg.n)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "neighbour mask stride below particle count");
    const size_t tasks = (sg.n + kThreads - 1) / kThreads;
    if (tasks > 0x7fffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "too many particles for the neighbour sweep");
    PBF_LAUNCH((k_sweep<Pass, kConsume>), static_cast<unsigned>(tasks), kThreads, stream, g, pa);
}

}  // namespace

namespace pbfint {

size_t sweep_tile_count(pbf_int3 dims) {
    const size_t chunks = (static_cast<size_t>(dims.x) + kTileW - 1) / kTileW;
    return chunks * static_cast<size_t>(dims.y) * static_cast<size_t>(dims.z);
}

size_t sweep_mask_words(size_t stride) {
    return static_cast<size_t>(kSegs * kRowWords) * stride;
}

pbf_status sweep_lambda(const SweepGrid& g, const pbf_float4* p, float* lambda, uint32_t* status,
                        float mass, float rho0, float eps, void* stream) {
    LambdaPass::Args a;
    a.p = p;
    a.lambda = lambda;
    a.status = status;
    a.mass = mass;
    a.rho0 = rho0;
    a.eps = eps;
    a.gs = mass / rho0;
    return launch_sweep<LambdaPass, false>(g, a, stream);
}

pbf_status sweep_delta_collide(const SweepGrid& g, const pbf_float4* p, const float* lambda,
                               pbf_float4* pout, uint32_t* status, float mass, float rho0,
                               float sk, int sn, float sdq, const pbfdev::Colliders& col,
                               float radius, void* stream) {
    DeltaPass::Args a;
    a.p = p;
    a.lambda = lambda;
    a.pout = pout;
    a.status = status;
    a.sk = sk;
    a.sn = sn;
    a.sdq = sdq;
    a.scale = mass / rho0;
    a.col = col;
    a.radius = radius;
    return launch_sweep<DeltaPass, true>(g, a, stream);
}

pbf_status sweep_update_xsph(const SweepGrid& g, const pbf_float4* x, const pbf_float4* p,
                             const pbf_float4* vold, pbf_float4* vout, uint32_t* status,
                             float visc, float invDt, void* stream) {
    XsphPass::Args a;
    a.x = x;
    a.p = p;
    a.vold = vold;
    a.vout = vout;
    a.status = status;
    a.visc = visc;
    a.invDt = invDt;
    return launch_sweep<XsphPass, false>(g, a, stream);
}

pbf_status sweep_vorticity(const SweepGrid& g, const pbf_float4* x, const pbf_float4* v,
                           pbf_float4* omega, uint32_t* status, bool reuseMasks, void* stream) {
    VorticityPass::Args a;
    a.p = x;
    a.v = v;
    a.omega = omega;
    a.status = status;
    return reuseMasks ? launch_sweep<VorticityPass, true>(g, a, stream)
                      : launch_sweep<VorticityPass, false>(g, a, stream);
}

pbf_status sweep_confine_resolve(const SweepGrid& g, const pbf_float4* x, const pbf_float4* omega,
                                 const pbf_float4* vcur, const pbf_float4* vsnap,
                                 pbf_float4* vout, uint32_t* status, float dt, float strength,
                                 const pbfdev::Colliders& col, float radius, float e, float f,
                                 void* stream) {
    ConfinePass::Args a;
    a.x = x;
    a.omega = omega;
    a.vcur = vcur;
    a.vsnap = vsnap;
    a.vout = vout;
    a.status = status;
    a.dt = dt;
    a.strength = strength;
    a.col = col;
    a.radius = radius;
    a.restitution = e;
    a.friction = f;
    return launch_sweep<ConfinePass, true>(g, a, stream);
}

}  // namespace pbfint
