// Neighbour lists of the solver step by LDS-tiled cell sweep (gfx950).
//
// Builds exactly the list findNeighbors builds (src/pbf/neighbors.cu:3-91):
// for particle i, every other particle j of the 3x3x3 cells around i's cell
// with d2 <= h*h (FP32, no FMA), in z, y, x cell order and sorted order
// within a cell (at most maxN entries; the count stays whole and raises the
// overflow bit above maxN).  The reference stores it slot-major at
// nbr[q * stride + i]; this build stores it in groups of four (below).
// The reference reads every candidate through global memory per particle;
// here the candidates are staged in LDS per tile:
//
//   * a tile is up to 32 consecutive cells of one grid row (fixed y, z).
//     Particles are sorted by cell key (x fastest), so the cells
//     x0-1 .. x1+1 of any row are one contiguous range of the working set:
//     a tile's neighbourhood is 9 such row segments (dz, dy in -1..1);
//   * sorted keys visit tiles in increasing order; workgroup b takes the
//     tiles whose first particle lies in [256 b, 256 b + 256), so no empty
//     tile is visited and no list of tiles is built;
//   * the 9 segments are staged one z-plane (3 rows) at a time, coordinates
//     in separate arrays, with coalesced loads;
//   * each thread owns one particle of the tile.  For every row it tests the
//     candidates of cells cx-1 .. cx+1 in windows of 32 (branch-free, four
//     loads in flight) and appends the set bits, lowest first, to its list;
//   * lists are stored in groups of four entries, group k of particle i at
//     nbr[k * stride + i] (16 bytes), so one store carries four entries and
//     the list kernels read four entries with one load.  Slots of the last
//     group past the count repeat its last entry, so every slot of a
//     written group is a valid index;
//   * a workgroup whose range starts more than kMaxTilePasses tiles (sparse
//     regions: spray, surface sheets) gives each of their particles a lane
//     of its own that reads its 27 cells from global memory instead of
//     sweeping the tiles one after another.
//
// Requires a 3x3x3 stencil (cell size >= the list radius), which the
// solver's grids guarantee (src/pbf/pbf_solver.cu:81-82 for the reference's
// cells of size h).
//
// Skin lists (the solver's one build per substep) use the same sweep with
// radius h * (1 + skin) over cells of that size, every candidate taken at
// the container point of its position.
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
constexpr int kMaxTilePasses = 4;      // more tiles per workgroup take the lane path
constexpr int kStageUnroll = 3;        // staging loads in flight per thread

struct Grid {
    const pbf_float4* p;
    const int* cellStart;
    const int* cellEnd;
    const uint32_t* sortedKeys;
    const uint32_t* tileEnd;
    uint4* nbr;  // grouped lists
    int* cnt;
    uint32_t* status;
    size_t n;
    size_t stride;
    pbf_int3 dims;
    int chunks;  // tiles per row = ceil(dims.x / kTileW)
    float h2;
    int maxNeighbors;
    pbfint::SceneTable st;  // a batched step's scenes
    // Skin lists: candidates are taken at their container points, and a
    // list past maxNeighbors marks the lists stale instead of overflowing.
    bool skin;
    pbf_container box;
    float radius;
    uint32_t fullBit;
};

// The grid a tile or particle belongs to: the one grid of a single-scene
// step, or its scene's grid in a batched step (keys and tiles offset into
// the global ranges).
struct Geom {
    pbf_int3 dims;
    uint32_t cellOff;
    uint32_t tileOff;
    int chunks;
    pbf_container box;  // the scene's container (skin lists)
};

template <bool kBatched>
__device__ __forceinline__ Geom geom_of(const Grid& g, size_t slot) {
    if (!kBatched) return Geom{g.dims, 0u, 0u, g.chunks, g.box};
    const uint32_t scene = g.st.sceneOf[slot];
    const pbfint::SceneGeom& sg = g.st.geom[scene];
    return Geom{sg.dims, sg.cellOff, sg.tileOff, static_cast<int>(sg.chunks),
                g.skin ? g.st.containers[scene] : pbf_container{}};
}

// The position a candidate is tested at: its own, or for skin lists its
// container point (the keys of a skin build are those of the same points).
template <bool kSkin>
__device__ __forceinline__ float3 test_point(const pbf_container& box, float radius, float4 p) {
    if (kSkin) return container_point(p, box, radius);
    return make_float3(p.x, p.y, p.z);
}

struct Meta {
    int segFirst[kSegs];  // working-set index of the segment's first particle
    int segBase[kSegs];   // LDS slot of that particle within its plane
    int segCells[kSegs];  // cells of the segment (0: empty or outside the grid)
    int planeCount[3];
    int cellOff[kSegs][kSegCells + 1];  // exclusive prefix of per-cell counts
    int stop;
    int waveTiles[kThreads / 64];
    int tiles[kThreads];      // tiles starting in this workgroup's particle range
    int tileFirst[kThreads];  // their first particles
};

// Staged coordinates of one z-plane as packed 12-byte (x, y, z) records:
// the twelve loads of a four-candidate step share one address register with
// immediate offsets, and lanes ~8 records apart (neighbouring cells) land
// 24 banks apart, where 16-byte records would share one bank group.
struct Plane {
    float xyz[3 * (kCap + kPad)];
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8).  Give every XCD one contiguous run of particle ranges so
// neighbouring tiles, which share halo rows, share that XCD's L2.
__device__ __forceinline__ int xcd_task(int b, int nb) {
    return static_cast<int>(xcd_block(static_cast<unsigned>(b), static_cast<unsigned>(nb)));
}

// Workgroup barrier for LDS hand-offs only.  No thread of the sweep reads
// global memory that another thread of the same launch writes, so the
// barrier orders LDS alone and leaves global loads in flight across it
// (__syncthreads would wait for every outstanding global load first).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t tile_of(uint32_t key, const Geom& gm) {
    const uint32_t k = key - gm.cellOff;
    const uint32_t gx = static_cast<uint32_t>(gm.dims.x);
    return gm.tileOff + (k / gx) * static_cast<uint32_t>(gm.chunks) +
           (k % gx) / static_cast<uint32_t>(kTileW);
}

// Candidates of one row segment: staged in LDS, or read from the working
// set when the plane did not fit.  Index q is relative to the segment's
// first particle; position() may be asked for up to 3 records past the row
// (padding in LDS, clamped in global memory).
struct LdsRow {
    const Plane* pl;
    int base;
    __device__ float3 position(int q) const {
        const float* r = pl->xyz + 3 * (base + q);
        return make_float3(r[0], r[1], r[2]);
    }
};

template <bool kSkin>
struct GlobalRow {
    const pbf_float4* p;
    size_t first;
    int last;
    pbf_container box;
    float radius;
    __device__ float3 position(int q) const {
        return test_point<kSkin>(box, radius, ld4(p, first + static_cast<size_t>(min(q, last))));
    }
};

// mask = 2 * mask + (d[u] <= h2) for u = 0, 1, 2, 3 in turn: one compare
// into a scalar pair and one add-with-carry per candidate (the compiler's
// select, shift and or take three).  The four compares come first, so each
// carry is read three instructions after it is written, past the two wait
// states a VALU-written scalar needs before a VALU reads it.  The compare is
// h2 >= d2, false for NaN exactly as d2 <= h2.
__device__ __forceinline__ uint32_t push4(uint32_t mask, float h2, float d0, float d1, float d2,
                                          float d3) {
    uint64_t c0, c1, c2, c3, co;
    asm("v_cmp_ge_f32_e64 %1, %6, %7\n\t"
        "v_cmp_ge_f32_e64 %2, %6, %8\n\t"
        "v_cmp_ge_f32_e64 %3, %6, %9\n\t"
        "v_cmp_ge_f32_e64 %4, %6, %10\n\t"
        "v_addc_co_u32_e64 %0, %5, %0, %0, %1\n\t"
        "v_addc_co_u32_e64 %0, %5, %0, %0, %2\n\t"
        "v_addc_co_u32_e64 %0, %5, %0, %0, %3\n\t"
        "v_addc_co_u32_e64 %0, %5, %0, %0, %4"
        : "+v"(mask), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3), "=&s"(co)
        : "s"(h2), "v"(d0), "v"(d1), "v"(d2), "v"(d3));
    return mask;
}

// A particle's list under construction.  buf holds the open group of four,
// newest entry in .w; count is the number of neighbours found (it may pass
// maxNeighbors), stored the number kept (at most maxNeighbors).
struct ListState {
    uint4 buf;
    int count;
    int stored;
};

// Appends the neighbours of particle gi among candidates [lo, hi) of one
// row (first = the row's first working-set index) to gi's list, in order.
template <class Src>
__device__ __forceinline__ void sweep_row(const Grid& g, const Src& src, float3 ip, int lo, int hi,
                                          int selfq, size_t gi, size_t first, ListState& st) {
    for (int c = lo; c < hi; c += 32) {
        const int w = min(32, hi - c);
        const int w4 = (w + 3) & ~3;
        uint32_t mask = 0u;
        for (int b = 0; b < w; b += 4) {
            float3 j[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) j[u] = src.position(c + b + u);
            float d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                d[u] = dist2_exact(ip.x - j[u].x, ip.y - j[u].y, ip.z - j[u].z);
            mask = push4(mask, g.h2, d[0], d[1], d[2], d[3]);
        }
        // Candidate b sits at bit w4 - 1 - b; reverse so that it sits at b.
        mask = __builtin_bitreverse32(mask) >> (32 - w4);
        if (w < 32) mask &= (1u << w) - 1u;
        if (selfq >= c && selfq < c + 32) mask &= ~(1u << (selfq - c));
        if (mask == 0u) continue;
        // Entries past maxNeighbors are counted, not stored: keep the
        // lowest free-slot-many bits (rare), so the loop below keeps all.
        const int hits = __builtin_popcount(mask);
        const int room = g.maxNeighbors - st.stored;
        st.count += hits;
        if (hits > room) {
            uint32_t kept = 0u;
            for (int k = 0; k < room; ++k) {
                const uint32_t low = mask & (0u - mask);
                kept |= low;
                mask ^= low;
            }
            mask = kept;
        }
        uint4* out = g.nbr + static_cast<size_t>(st.stored >> 2) * g.stride + gi;
        const uint32_t base = static_cast<uint32_t>(first) + static_cast<uint32_t>(c);
        while (mask != 0u) {
            const uint32_t e = base + static_cast<uint32_t>(__builtin_ctz(mask));
            mask &= mask - 1u;
            st.buf = make_uint4(st.buf.y, st.buf.z, st.buf.w, e);
            if ((++st.stored & 3) == 0) {
                *out = st.buf;
                out += g.stride;
            }
        }
    }
}

// Stores the open group, its free slots repeating the last entry, and the
// count.
__device__ __forceinline__ void finish_particle(const Grid& g, size_t gi, const ListState& st) {
    const int k = st.stored & 3;
    if (k != 0) {
        const uint4 b = st.buf;
        const uint4 last = k == 1   ? make_uint4(b.w, b.w, b.w, b.w)
                           : k == 2 ? make_uint4(b.z, b.w, b.w, b.w)
                                    : make_uint4(b.y, b.z, b.w, b.w);
        g.nbr[static_cast<size_t>(st.stored >> 2) * g.stride + gi] = last;
    }
    g.cnt[gi] = st.count;
    if (st.count > g.maxNeighbors) atomicOr(g.status, g.fullBit);
}

// One tile: segment geometry, then the tile's particles against the three
// z-planes of its halo.  Every branch that reaches a barrier is uniform
// over the workgroup (it depends on LDS values written before a barrier).
template <bool kSkin>
__device__ __forceinline__ void sweep_tile(const Grid& g, const Geom& gm, Meta& m, Plane& lds,
                                           int tile) {
    const int local = tile - static_cast<int>(gm.tileOff);
    const int row = local / gm.chunks;
    const int chunk = local - row * gm.chunks;
    const int y = row % gm.dims.y;
    const int z = row / gm.dims.y;
    const int x0 = chunk * kTileW;
    const int x1 = min(x0 + kTileW, gm.dims.x) - 1;
    const int xs = max(x0 - 1, 0);
    const int xe = min(x1 + 1, gm.dims.x - 1);
    const int ncell = xe - xs + 1;
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // Segment geometry: one wave per row, one lane per cell.
    for (int s = wave; s < kSegs; s += kThreads / 64) {
        const int yy = y + (s % 3) - 1;
        const int zz = z + (s / 3) - 1;
        const bool rowOk = yy >= 0 && yy < gm.dims.y && zz >= 0 && zz < gm.dims.z;
        int cnt = 0;
        int st = 0x7fffffff;
        if (rowOk && lane < ncell) {
            const uint32_t key = gm.cellOff + key_of(xs + lane, yy, zz, gm.dims);
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
        float3 ip = make_float3(0.0f, 0.0f, 0.0f);
        int mc = 0;
        ListState st{make_uint4(0u, 0u, 0u, 0u), 0, 0};
        if (active) {
            ip = test_point<kSkin>(gm.box, g.radius, ld4(g.p, gi));
            mc = static_cast<int>((g.sortedKeys[gi] - gm.cellOff) %
                                  static_cast<uint32_t>(gm.dims.x)) - xs;
        }
        for (int pl = 0; pl < 3; ++pl) {
            const int pc = m.planeCount[pl];
            if (pc == 0) continue;
            const bool staged = pc <= kCap;
            if (staged) {
                lds_barrier();  // the previous plane's readers are done
                // The plane's three segments are consecutive in LDS: record q
                // of [0, pc) is row 0 below segBase of row 1, row 1 below
                // segBase of row 2, else row 2 (an empty row has no q).  A
                // thread issues kStageUnroll loads before the first store.
                const int s0 = 3 * pl;
                const int b1 = m.segBase[s0 + 1];
                const int b2 = m.segBase[s0 + 2];
                const size_t f0 = static_cast<size_t>(m.segFirst[s0]);
                const size_t f1 = static_cast<size_t>(m.segFirst[s0 + 1]) - b1;
                const size_t f2 = static_cast<size_t>(m.segFirst[s0 + 2]) - b2;
                for (int q0 = tid; q0 < pc; q0 += kStageUnroll * kThreads) {
                    float4 v[kStageUnroll];
#pragma unroll
                    for (int u = 0; u < kStageUnroll; ++u) {
                        const int q = q0 + u * kThreads;
                        if (q < pc) v[u] = ld4(g.p, (q < b1 ? f0 : (q < b2 ? f1 : f2)) + q);
                    }
#pragma unroll
                    for (int u = 0; u < kStageUnroll; ++u) {
                        const int q = q0 + u * kThreads;
                        if (q < pc) {
                            const float3 a = test_point<kSkin>(gm.box, g.radius, v[u]);
                            float* r = lds.xyz + 3 * q;
                            r[0] = a.x;
                            r[1] = a.y;
                            r[2] = a.z;
                        }
                    }
                }
                lds_barrier();
            }
            if (!active) continue;
            for (int rr = 0; rr < 3; ++rr) {
                const int s = 3 * pl + rr;
                const int nc = m.segCells[s];
                if (nc == 0) continue;
                const int lo = m.cellOff[s][max(mc - 1, 0)];
                const int hi = m.cellOff[s][min(mc + 1, nc - 1) + 1];
                if (hi <= lo) continue;
                const int selfq = s == 4 ? r : -1;
                const size_t first = static_cast<size_t>(m.segFirst[s]);
                if (staged) {
                    const LdsRow src{&lds, m.segBase[s]};
                    sweep_row(g, src, ip, lo, hi, selfq, gi, first, st);
                } else {
                    const GlobalRow<kSkin> src{g.p, first, m.cellOff[s][nc] - 1, gm.box,
                                               g.radius};
                    sweep_row(g, src, ip, lo, hi, selfq, gi, first, st);
                }
            }
        }
        if (active) finish_particle(g, gi, st);
    }
    lds_barrier();  // LDS is reused by the next tile
}

// One particle on one lane, candidates read from global memory: the row
// range of every (dz, dy) is the run of particles from the first non-empty
// cell of cx-1 .. cx+1 to the last, as in the tile form.
template <bool kBatched, bool kSkin>
__device__ __forceinline__ void sweep_lane(const Grid& g, size_t gi) {
    const Geom gm = geom_of<kBatched>(g, gi);
    const uint32_t key = g.sortedKeys[gi] - gm.cellOff;
    const uint32_t gx = static_cast<uint32_t>(gm.dims.x);
    const int cx = static_cast<int>(key % gx);
    const uint32_t rowId = key / gx;
    const int cy = static_cast<int>(rowId % static_cast<uint32_t>(gm.dims.y));
    const int cz = static_cast<int>(rowId / static_cast<uint32_t>(gm.dims.y));
    const float3 ip = test_point<kSkin>(gm.box, g.radius, ld4(g.p, gi));
    ListState st{make_uint4(0u, 0u, 0u, 0u), 0, 0};
    const int xa = max(cx - 1, 0);
    const int xb = min(cx + 1, gm.dims.x - 1);
    for (int s = 0; s < kSegs; ++s) {
        const int yy = cy + (s % 3) - 1;
        const int zz = cz + (s / 3) - 1;
        if (yy < 0 || yy >= gm.dims.y || zz < 0 || zz >= gm.dims.z) continue;
        int rowStart = -1, rowEnd = -1;
        for (int x = xa; x <= xb; ++x) {
            const uint32_t k = gm.cellOff + key_of(x, yy, zz, gm.dims);
            const int a = g.cellStart[k];
            if (a == -1) continue;
            if (rowStart < 0) rowStart = a;
            rowEnd = g.cellEnd[k];
        }
        if (rowStart < 0) continue;
        const int len = rowEnd - rowStart;
        const int selfq = s == 4 ? static_cast<int>(gi - static_cast<size_t>(rowStart)) : -1;
        const GlobalRow<kSkin> src{g.p, static_cast<size_t>(rowStart), len - 1, gm.box,
                                   g.radius};
        sweep_row(g, src, ip, 0, len, selfq, gi, static_cast<size_t>(rowStart), st);
    }
    finish_particle(g, gi, st);
}

// One workgroup per 256 sorted particles: it sweeps every tile whose first
// particle falls in its range (none, one, or several small ones).  Seven
// waves per SIMD (72 VGPRs, three spilled on cold paths) rather than the six
// that 77 VGPRs allow: the sweep waits on LDS and the list stores most of
// the time, and the seventh wave hides part of it.
template <bool kBatched, bool kSkin>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_build_lists(Grid g) {
    __shared__ Meta m;
    __shared__ Plane lds;

    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const size_t i = static_cast<size_t>(xcd_task(static_cast<int>(blockIdx.x),
                                                  static_cast<int>(gridDim.x))) *
                         kThreads + static_cast<size_t>(tid);
    bool starts = false;
    uint32_t t = 0;
    if (i < g.n) {
        t = tile_of(g.sortedKeys[i], geom_of<kBatched>(g, i));
        starts = i == 0 || tile_of(g.sortedKeys[i - 1], geom_of<kBatched>(g, i - 1)) != t;
    }
    // Tiles in particle order: wave counts, then ranks within the wave.
    const uint64_t ballot = __ballot(starts);
    if (lane == 0) m.waveTiles[wave] = __popcll(ballot);
    // A raised status word stops the step; read once, through LDS, so the
    // whole workgroup takes the decision together.
    if (tid == 0) m.stop = *g.status != 0u ? 1 : 0;
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
            sweep_tile<kSkin>(g, geom_of<kBatched>(g, static_cast<size_t>(m.tileFirst[k])), m,
                              lds, m.tiles[k]);
    } else {
        // The particles of every tile that starts here, one lane each.
        const size_t first = static_cast<size_t>(m.tileFirst[0]);
        const size_t end = g.tileEnd[m.tiles[total - 1]];
        for (size_t q = first + static_cast<size_t>(tid); q < end; q += kThreads)
            sweep_lane<kBatched, kSkin>(g, q);
    }
}

}  // namespace

namespace pbfint {

size_t sweep_tile_count(pbf_int3 dims) {
    const size_t chunks = (static_cast<size_t>(dims.x) + kTileW - 1) / kTileW;
    return chunks * static_cast<size_t>(dims.y) * static_cast<size_t>(dims.z);
}

pbf_status sweep_build_lists(const SweepGrid& sg, const pbf_float4* p, uint32_t* nbr, int* cnt,
                             size_t stride, uint32_t* status, void* stream) {
    if (sg.n == 0) return PBF_OK;
    if (stride < sg.n)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "neighbour list stride below particle count");
    if (sg.maxNeighbors < 0 || sg.maxNeighbors % 4 != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "grouped lists need a multiple of 4 slots");
    if (reinterpret_cast<uintptr_t>(nbr) % alignof(uint4) != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "grouped lists need 16-byte alignment");
    Grid g;
    g.p = p;
    g.cellStart = sg.cellStart;
    g.cellEnd = sg.cellEnd;
    g.sortedKeys = sg.sortedKeys;
    g.tileEnd = sg.tileEnd;
    g.nbr = reinterpret_cast<uint4*>(nbr);
    g.cnt = cnt;
    g.status = status;
    g.n = sg.n;
    g.stride = stride;
    g.dims = sg.dims;
    g.chunks = (sg.dims.x + kTileW - 1) / kTileW;
    g.h2 = sg.h * sg.h;
    g.maxNeighbors = sg.maxNeighbors;
    g.st = sg.st;
    g.skin = sg.skin;
    g.box = sg.box != nullptr ? *sg.box : pbf_container{};
    g.radius = sg.radius;
    g.fullBit = sg.skin ? PBF_STATUS_LIST_STALE : PBF_STATUS_NEIGHBOR_OVERFLOW;
    if (sg.skin && sg.st.geom == nullptr && sg.box == nullptr)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "skin lists need the container");
    const size_t tiles = sg.st.geom != nullptr ? 0 : sweep_tile_count(sg.dims);
    if (tiles > 0x7fffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "grid has too many rows for the neighbour sweep");
    const size_t tasks = (sg.n + kThreads - 1) / kThreads;
    if (tasks > 0x7fffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "too many particles for the neighbour sweep");
    const unsigned nb = static_cast<unsigned>(tasks);
    if (sg.st.geom != nullptr && sg.skin)
        PBF_LAUNCH((k_build_lists<true, true>), nb, kThreads, stream, g);
    if (sg.st.geom != nullptr) PBF_LAUNCH((k_build_lists<true, false>), nb, kThreads, stream, g);
    if (sg.skin) PBF_LAUNCH((k_build_lists<false, true>), nb, kThreads, stream, g);
    PBF_LAUNCH((k_build_lists<false, false>), nb, kThreads, stream, g);
}

}  // namespace pbfint
