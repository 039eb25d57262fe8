// Neighbour lists of the solver step by LDS-staged cell sweep (gfx950).
//
// Builds exactly the list findNeighbors builds (src/pbf/neighbors.cu:3-91):
// for particle i, every other particle j of the 3x3x3 cells around i's cell
// with d2 <= h*h (FP32, no FMA), in z, y, x cell order and sorted order
// within a cell (at most maxN entries; the count stays whole and raises the
// overflow bit above maxN).  The reference stores it slot-major at
// nbr[q * stride + i]; this build stores it in groups of four (below).
// The reference reads every candidate through global memory per particle;
// here each wave stages its particles' candidates in LDS:
//
//   * a wave takes 64 consecutive particles of the sorted working set, one
//     per lane, and works on its own (no workgroup barrier): every lane is
//     busy whatever the cell occupancy;
//   * particles are sorted by cell key (x fastest), so the wave's particles
//     of one grid row (a "piece") occupy cells xa .. xb of that row, and for
//     every (dz, dy) the cells xa-1 .. xb+1 of the row at (y+dy, z+dz) are
//     one contiguous run of the working set (a segment).  A wave holds at
//     most two pieces of at most kMaxWidth cells each; any other wave (spray,
//     sparse surface rows) gives every lane a sweep of its own through
//     global memory;
//   * per (piece, dz, dy) sixteen lanes read the segment's cell ranges and
//     scan them into per-cell offsets in LDS;
//   * row by row (dz, dy), the wave stages both pieces' segments in LDS as
//     packed (x, y, z) and each lane tests the candidates of its cells
//     cx-1 .. cx+1 in windows of 32 (four compares per step folded into a
//     32-bit hit mask, candidates in order from the top bit down);
//   * the set bits of a window are appended four per pass through a per-lane
//     LDS column that holds the open group of four: the pass writes four
//     entries after the open ones, and a filled group leaves as one 16-byte
//     store.  The wave runs as many passes as its busiest lane needs quarter
//     hits, not hits;
//   * lists are stored in groups of four entries, group k of particle i at
//     nbr[k * stride + i] (16 bytes), so one store carries four entries and
//     the list kernels read four entries with one load.  Slots of the last
//     group past the count repeat its last entry, so every slot of a
//     written group is a valid index;
//   * the post-solve build of a step with XSPH (kXsph) also stages each
//     candidate's velocity (p* - x) / dt beside its position, adds each
//     appended neighbour's applyXsphViscosity term
//     (src/pbf/pbf_solver_kernels.cu:213-264) to the lane's sum in list
//     order, and stores the resulting velocity: the step then needs no
//     separate XSPH pass, and the velocities are bitwise the separate
//     pass's.
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

constexpr int kWaves = 4;                  // independent waves per workgroup
constexpr int kThreads = 64 * kWaves;
constexpr int kSegCells = 16;              // cells of a segment: one scan group
constexpr int kMaxWidth = kSegCells - 2;   // cells a piece may span
constexpr int kPieces = 2;                 // grid rows a staged wave may span
constexpr int kSegs = 9;                   // (dz, dy) rows, dz-major
constexpr int kStage = 160;                // staged records per row (both pieces)
constexpr int kPad = 4;                    // test reads run up to 3 records past a window
constexpr int kStageUnroll = 3;            // staging loads in flight per lane
constexpr int kBufRows = 8;                // per-lane append column (see append)

struct Grid {
    const pbf_float4* p;
    const int* cellStart;
    const int* cellEnd;
    const uint32_t* sortedKeys;
    uint4* nbr;  // grouped lists
    int* cnt;
    uint32_t* status;
    size_t n;
    size_t stride;
    pbf_int3 dims;
    float h2;
    int maxNeighbors;
    pbfint::SceneTable st;  // a batched step's scenes
    // Skin lists: candidates are taken at their container points, and a
    // list past maxNeighbors marks the lists stale instead of overflowing.
    bool skin;
    pbf_container box;
    float radius;
    uint32_t fullBit;
    // XSPH of the post-solve build (kXsph): start positions, velocities in
    // (.w kept) and out, the tape's copy of the output (or null), the
    // viscosity, 1 / dt and h.
    const pbf_float4* xs;
    const pbf_float4* vold;
    pbf_float4* vout;
    pbf_float4* tapeOut;
    float visc;
    float inv_dt;
    float h;
};

// The grid a particle belongs to: the one grid of a single-scene step, or
// its scene's grid in a batched step (keys offset into the global range).
struct Geom {
    pbf_int3 dims;
    uint32_t cellOff;
    pbf_container box;  // the scene's container (skin lists)
};

template <bool kBatched>
__device__ __forceinline__ Geom geom_of(const Grid& g, size_t slot) {
    if (!kBatched) return Geom{g.dims, 0u, g.box};
    const uint32_t scene = g.st.sceneOf[slot];
    const pbfint::SceneGeom& sg = g.st.geom[scene];
    return Geom{sg.dims, sg.cellOff, g.skin ? g.st.containers[scene] : pbf_container{}};
}

// The position a candidate is tested at: its own, or for skin lists its
// container point (the keys of a skin build are those of the same points).
template <bool kSkin>
__device__ __forceinline__ float3 test_point(const pbf_container& box, float radius, float4 p) {
    if (kSkin) return container_point(p, box, radius);
    return make_float3(p.x, p.y, p.z);
}

// One wave's LDS: the staged row, the segment offsets, and the lanes' append
// columns.  Coordinates are packed 12-byte (x, y, z) records: the twelve
// loads of a four-candidate step share one address register with immediate
// offsets, and lanes ~8 records apart (neighbouring cells) land 24 banks
// apart, where 16-byte records would share one bank group.
// The XSPH build also stages each candidate's velocity (v = (p* - x) / dt)
// beside its position.
template <bool kXsph>
struct WaveLds {
    float xyz[3 * (kStage + kPad)];
    float vel[kXsph ? 3 * (kStage + kPad) : 1];
    uint16_t off[kPieces][kSegs][kSegCells + 1];  // exclusive prefix of per-cell counts
    int first[kPieces][kSegs];  // working-set index of the segment's start
    uint32_t buf[kBufRows][64];  // row r of lane l at buf[r][l]
};

// A neighbour's XSPH inputs: its p* and velocity.
struct XRec {
    float3 p, v;
};
// The lane's XSPH sum: its own p* and velocity, the sum so far.
struct XsphSum {
    float3 pi, vi, c;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8).  Give every XCD one contiguous run of particle ranges so
// neighbouring waves, which share halo rows, share that XCD's L2.
__device__ __forceinline__ int xcd_task(int b, int nb) {
    return static_cast<int>(xcd_block(static_cast<unsigned>(b), static_cast<unsigned>(nb)));
}

// Orders one wave's LDS accesses for the compiler.  A wave's LDS operations
// execute in order, so what one lane writes is visible to every lane's later
// reads without a workgroup barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Candidates of one segment: staged in LDS, or read from the working set
// when the row did not fit.  Index q is relative to the segment's first
// particle; position() may be asked for up to 3 records past a window
// (padding in LDS, clamped in global memory).
struct LdsRow {
    const float* xyz;
    const float* vel;
    int base;
    __device__ float3 position(int q) const {
        const float* r = xyz + 3 * (base + q);
        return make_float3(r[0], r[1], r[2]);
    }
    __device__ XRec record(int q, uint32_t) const {
        const float* v = vel + 3 * (base + q);
        return XRec{position(q), make_float3(v[0], v[1], v[2])};
    }
};

template <bool kSkin>
struct GlobalRow {
    const pbf_float4* p;
    size_t first;
    int last;
    pbf_container box;
    float radius;
    const pbf_float4* xs;  // XSPH: start positions
    float inv_dt;
    __device__ float3 position(int q) const {
        return test_point<kSkin>(box, radius, ld4(p, first + static_cast<size_t>(min(q, last))));
    }
    __device__ XRec record(int, uint32_t j) const {
        const float4 pj = ld4(p, j), xj = ld4(xs, j);
        return XRec{make_float3(pj.x, pj.y, pj.z),
                    make_float3(stored_velocity(pj.x, xj.x, inv_dt),
                                stored_velocity(pj.y, xj.y, inv_dt),
                                stored_velocity(pj.z, xj.z, inv_dt))};
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

// A particle's list under construction: the next group's address, the
// neighbours found (may pass maxNeighbors), the entries kept (at most
// maxNeighbors), and how many of them wait in the open group (rows 0 ..
// open - 1 of the lane's append column).
struct ListState {
    uint4* out;
    int count;
    int stored;
    int open;
};

// Appends the set bits of a window's mask (candidate b at bit 31 - (lz of
// b), read from the top down, entry base + lz) to the list.
//
// Each pass takes the next four set bits (the entries past the lane's last
// are garbage and land beyond its open group) and writes them to rows
// open .. open + 3 of the lane's column; when the open group (rows 0 .. 3)
// fills, it is stored and the entries past it move down four rows.
//
// The XSPH build (kXsph) adds each kept entry's XSPH term, in list order,
// from the candidate's record (src.record(q), q = qbase + lz its index in
// the segment).
template <bool kXsph, class Src>
__device__ __forceinline__ void append(const Grid& g, uint32_t mask, uint32_t base, int qbase,
                                       ListState& st, WaveLds<kXsph>& L, int lane,
                                       const Src& src, XsphSum& xs, const Sph& sph) {
    int hits = __builtin_popcount(mask);
    st.count += hits;
    const int room = g.maxNeighbors - st.stored;
    if (hits > room) {
        // Entries past maxNeighbors are counted, not stored: keep the first
        // room-many candidates (the highest bits).
        for (int k = hits - room; k > 0; --k) mask &= mask - 1u;
        hits = room;
    }
    for (int rem = hits; rem > 0; rem -= 4) {
        uint32_t e[4];
        int lzs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int lz = __builtin_clzg(mask, -1);
            lzs[u] = lz;
            e[u] = base + static_cast<uint32_t>(lz);
            mask &= 0x7fffffffu >> (lz & 31);
        }
        if (kXsph) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < rem) {
                    const XRec r = src.record(qbase + lzs[u], e[u]);
                    xsph_term(xs.pi, xs.vi, r.p, r.v, sph, xs.c);
                }
        }
        uint32_t* col = &L.buf[st.open][lane];
#pragma unroll
        for (int u = 0; u < 4; ++u) col[64 * u] = e[u];
        const int taken = min(rem, 4);
        st.stored += taken;
        const int open = st.open + taken;
        if (open >= 4) {
            *st.out = make_uint4(L.buf[0][lane], L.buf[1][lane], L.buf[2][lane], L.buf[3][lane]);
            st.out += g.stride;
            // The entries past the stored group move down four rows; the
            // others go to row 7, free once the group is stored.
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = st.open + u - 4;
                L.buf[r >= 0 ? r : 7][lane] = e[u];
            }
            st.open = open - 4;
        } else {
            st.open = open;
        }
    }
}

// Appends the neighbours of the particle at test point ip among candidates
// [lo, hi) of one segment (first = the segment's first working-set index;
// selfq = the particle's own index in the segment, or -1) to its list.
template <bool kXsph, class Src>
__device__ __forceinline__ void sweep_row(const Grid& g, const Src& src, float3 ip, int lo, int hi,
                                          int selfq, size_t first, ListState& st,
                                          WaveLds<kXsph>& L, int lane, XsphSum& xs,
                                          const Sph& sph) {
    for (int c = lo; c < hi; c += 32) {
        const int w = min(32, hi - c);
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
        // Candidate b sits at bit w4 - 1 - b: drop the step's padding
        // candidates (b >= w, the lowest bits) and the particle itself.
        const int w4 = (w + 3) & ~3;
        mask &= ~0u << (w4 - w);
        const int sq = selfq - c;
        if (static_cast<unsigned>(sq) < static_cast<unsigned>(w)) mask &= ~(1u << (w4 - 1 - sq));
        if (mask == 0u) continue;
        append<kXsph>(g, mask, static_cast<uint32_t>(first) + static_cast<uint32_t>(c + w4 - 32),
                      c + w4 - 32, st, L, lane, src, xs, sph);
    }
}

// Stores the open group, its free slots repeating the last entry, and the
// count; the XSPH build also the velocity (NaN for an invalid count, as
// applyXsphViscosity writes, pbf_solver_kernels.cu:225-233).
template <bool kXsph>
__device__ __forceinline__ void finish_particle(const Grid& g, size_t gi, const ListState& st,
                                                const WaveLds<kXsph>& L, int lane,
                                                const XsphSum& xs, float vw) {
    if (st.open > 0) {
        const uint32_t a = L.buf[0][lane];
        const uint32_t b = st.open > 1 ? L.buf[1][lane] : a;
        const uint32_t c = st.open > 2 ? L.buf[2][lane] : b;
        *st.out = make_uint4(a, b, c, c);
    }
    g.cnt[gi] = st.count;
    if (st.count > g.maxNeighbors) atomicOr(g.status, g.fullBit);
    if (kXsph) {
        const float4 out =
            valid_count(st.count, g.maxNeighbors)
                ? make_float4(xs.vi.x + g.visc * xs.c.x, xs.vi.y + g.visc * xs.c.y,
                              xs.vi.z + g.visc * xs.c.z, vw)
                : make_float4(qnan(), qnan(), qnan(), vw);
        st4(g.vout, gi, out);
        if (g.tapeOut != nullptr) st4(g.tapeOut, gi, out);
    }
}

// One particle on one lane, candidates read from global memory: the row
// range of every (dz, dy) is the run of particles from the first non-empty
// cell of cx-1 .. cx+1 to the last.
template <bool kSkin, bool kXsph>
__device__ __forceinline__ void sweep_lane(const Grid& g, const Geom& gm, size_t gi, uint32_t key,
                                           ListState& st, WaveLds<kXsph>& L, int lane,
                                           XsphSum& xs, const Sph& sph) {
    const uint32_t gx = static_cast<uint32_t>(gm.dims.x);
    const int cx = static_cast<int>(key % gx);
    const uint32_t rowId = key / gx;
    const int cy = static_cast<int>(rowId % static_cast<uint32_t>(gm.dims.y));
    const int cz = static_cast<int>(rowId / static_cast<uint32_t>(gm.dims.y));
    const float3 ip = test_point<kSkin>(gm.box, g.radius, ld4(g.p, gi));
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
        const GlobalRow<kSkin> src{g.p, static_cast<size_t>(rowStart), len - 1, gm.box, g.radius,
                                   g.xs, g.inv_dt};
        sweep_row<kXsph>(g, src, ip, 0, len, selfq, static_cast<size_t>(rowStart), st, L, lane, xs,
                         sph);
    }
}

// One wave per 64 sorted particles (four independent waves per workgroup).
// Eight workgroups fit a CU's LDS (4.6 KB per wave), eight waves per SIMD:
// more waves hide more of the rows' staging loads (measured: 6, 7 and 8
// waves per SIMD give 1047, 992 and 966 us per launch on the 8M double dam
// break).  The XSPH build's staged velocities leave room for six.
template <bool kBatched, bool kSkin, bool kXsph>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kXsph ? 6 : 8, 8)))
k_build_lists(Grid g) {
    __shared__ WaveLds<kXsph> ldsAll[kWaves];
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int wv = static_cast<int>(threadIdx.x) >> 6;
    WaveLds<kXsph>& L = ldsAll[wv];
    const size_t w0 = (static_cast<size_t>(xcd_task(static_cast<int>(blockIdx.x),
                                                    static_cast<int>(gridDim.x))) *
                           kWaves + static_cast<size_t>(wv)) * 64;
    if (w0 >= g.n) return;
    // A raised status word stops the step.
    if (__builtin_amdgcn_readfirstlane(*g.status) != 0u) return;

    const size_t gi = w0 + static_cast<size_t>(lane);
    const bool active = gi < g.n;
    const size_t gc = active ? gi : g.n - 1;
    const int lastLane = static_cast<int>(min(g.n - 1 - w0, static_cast<size_t>(63)));
    const Geom gm = geom_of<kBatched>(g, gc);
    const uint32_t skey = g.sortedKeys[gc];
    const uint32_t key = skey - gm.cellOff;
    const uint32_t gx = static_cast<uint32_t>(gm.dims.x);
    const int cx = static_cast<int>(key % gx);
    const uint32_t rowId = key / gx;
    const int cy = static_cast<int>(rowId % static_cast<uint32_t>(gm.dims.y));
    const int cz = static_cast<int>(rowId / static_cast<uint32_t>(gm.dims.y));
    // The key of cell x = 0 of the particle's row names the row across scenes.
    const uint32_t rowKey = skey - static_cast<uint32_t>(cx);
    // XSPH: the lane's own p* and velocity (k_update_xsph's inputs).
    const Sph sph = sph_consts(g.h);
    XsphSum xv{};
    float vw = 0.0f;
    if (kXsph) {
        const float4 pp = ld4(g.p, gc), xx = ld4(g.xs, gc);
        xv.pi = make_float3(pp.x, pp.y, pp.z);
        xv.vi = make_float3(stored_velocity(pp.x, xx.x, g.inv_dt),
                            stored_velocity(pp.y, xx.y, g.inv_dt),
                            stored_velocity(pp.z, xx.z, g.inv_dt));
        vw = g.vold[gc].w;
    }

    // Pieces: runs of lanes in one grid row.
    const uint32_t prevRow = __shfl_up(rowKey, 1, 64);
    const uint64_t starts = __ballot(lane == 0 || prevRow != rowKey);
    const int pieces = __popcll(starts);
    const int split = pieces > 1 ? __builtin_ctzll(starts & ~1ull) : 64;
    const int first1 = min(split, 63);
    const int xa0 = __shfl(cx, 0, 64), xb0 = __shfl(cx, min(split - 1, lastLane), 64);
    const int xa1 = __shfl(cx, first1, 64), xb1 = __shfl(cx, lastLane, 64);
    const bool wide = xb0 - xa0 >= kMaxWidth || (pieces > 1 && xb1 - xa1 >= kMaxWidth);

    ListState st{g.nbr + gi, 0, 0, 0};
    if (pieces > kPieces || wide) {
        if (active) {
            sweep_lane<kSkin, kXsph>(g, gm, gi, key, st, L, lane, xv, sph);
            finish_particle<kXsph>(g, gi, st, L, lane, xv, vw);
        }
        return;
    }

    // Segment offsets: sixteen lanes per (piece, row), one cell each.  A
    // segment of more than 65,535 particles (offsets are 16-bit) sends the
    // wave to the per-lane sweep.
    const int pairs = pieces * kSegs;
    bool deep = false;
    for (int t0 = 0; t0 < pairs; t0 += 4) {
        const int t = t0 + (lane >> 4);
        const int k = lane & 15;
        const int p = t >= kSegs ? 1 : 0;
        const int s = t - p * kSegs;
        const int src = p != 0 ? first1 : 0;
        const int pcy = __shfl(cy, src, 64);
        const int pcz = __shfl(cz, src, 64);
        const Geom pg = kBatched ? geom_of<kBatched>(g, w0 + static_cast<size_t>(src)) : gm;
        const int xs = max((p != 0 ? xa1 : xa0) - 1, 0);
        const int xe = min((p != 0 ? xb1 : xb0) + 1, pg.dims.x - 1);
        const int yy = pcy + (s % 3) - 1;
        const int zz = pcz + (s / 3) - 1;
        int cnt = 0;
        int a = INT_MAX;
        if (t < pairs && xs + k <= xe && yy >= 0 && yy < pg.dims.y && zz >= 0 &&
            zz < pg.dims.z) {
            const uint32_t ck = pg.cellOff + key_of(xs + k, yy, zz, pg.dims);
            const int cs = g.cellStart[ck];
            if (cs != -1) {
                cnt = g.cellEnd[ck] - cs;
                a = cs;
            }
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int v = __shfl_up(incl, o, 16);
            if (k >= o) incl += v;
        }
        deep |= incl > 0xffff;
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) a = min(a, __shfl_xor(a, o, 16));
        if (t < pairs) {
            L.off[p][s][k + 1] = incl;
            if (k == 0) {
                L.off[p][s][0] = 0;
                L.first[p][s] = a == INT_MAX ? 0 : a;
            }
        }
    }
    wave_sync();
    if (__ballot(deep) != 0ull) {
        if (active) {
            sweep_lane<kSkin, kXsph>(g, gm, gi, key, st, L, lane, xv, sph);
            finish_particle<kXsph>(g, gi, st, L, lane, xv, vw);
        }
        return;
    }

    // The lane's piece, its segment cells and its cell among them.
    const int piece = lane >= split ? 1 : 0;
    const int xs = max((piece != 0 ? xa1 : xa0) - 1, 0);
    const int ncell = min((piece != 0 ? xb1 : xb0) + 1, gm.dims.x - 1) - xs + 1;
    const int mc = cx - xs;
    const int loCell = max(mc - 1, 0);
    const int hiCell = min(mc + 1, ncell - 1) + 1;
    const float3 ip = test_point<kSkin>(gm.box, g.radius, ld4(g.p, gc));
    const pbf_container box0 = kSkin ? geom_of<kBatched>(g, w0).box : pbf_container{};
    const pbf_container box1 =
        kSkin ? geom_of<kBatched>(g, w0 + static_cast<size_t>(first1)).box : pbf_container{};

    for (int s = 0; s < kSegs; ++s) {
        const int len0 = L.off[0][s][kSegCells];
        const int len1 = pieces > 1 ? L.off[1][s][kSegCells] : 0;
        const int total = len0 + len1;
        if (total == 0) continue;
        const bool staged = total <= kStage;
        if (staged) {
            // Segment 0 then segment 1, kStageUnroll loads in flight per lane.
            const long long f0 = L.first[0][s];
            const long long f1 = pieces > 1 ? static_cast<long long>(L.first[1][s]) - len0 : 0;
            for (int q0 = lane; q0 < total; q0 += 64 * kStageUnroll) {
                float4 v[kStageUnroll], x[kXsph ? kStageUnroll : 1];
#pragma unroll
                for (int u = 0; u < kStageUnroll; ++u) {
                    const int q = q0 + 64 * u;
                    if (q < total) {
                        const size_t at = static_cast<size_t>((q < len0 ? f0 : f1) + q);
                        v[u] = ld4(g.p, at);
                        if (kXsph) x[kXsph ? u : 0] = ld4(g.xs, at);
                    }
                }
#pragma unroll
                for (int u = 0; u < kStageUnroll; ++u) {
                    const int q = q0 + 64 * u;
                    if (q < total) {
                        const float3 r = test_point<kSkin>(q < len0 ? box0 : box1, g.radius, v[u]);
                        float* o = L.xyz + 3 * q;
                        o[0] = r.x;
                        o[1] = r.y;
                        o[2] = r.z;
                        if (kXsph) {
                            const float4 xq = x[kXsph ? u : 0];
                            float* w = L.vel + 3 * q;
                            w[0] = stored_velocity(v[u].x, xq.x, g.inv_dt);
                            w[1] = stored_velocity(v[u].y, xq.y, g.inv_dt);
                            w[2] = stored_velocity(v[u].z, xq.z, g.inv_dt);
                        }
                    }
                }
            }
            wave_sync();
        }
        if (active) {
            const int lo = L.off[piece][s][loCell];
            const int hi = L.off[piece][s][hiCell];
            if (hi > lo) {
                const size_t first = static_cast<size_t>(L.first[piece][s]);
                const int selfq = s == 4 ? static_cast<int>(gi - first) : -1;
                if (staged) {
                    const LdsRow src{L.xyz, L.vel, piece != 0 ? len0 : 0};
                    sweep_row<kXsph>(g, src, ip, lo, hi, selfq, first, st, L, lane, xv, sph);
                } else {
                    const GlobalRow<kSkin> src{g.p,      first,  (piece != 0 ? len1 : len0) - 1,
                                               gm.box,   g.radius, g.xs, g.inv_dt};
                    sweep_row<kXsph>(g, src, ip, lo, hi, selfq, first, st, L, lane, xv, sph);
                }
            }
        }
        wave_sync();  // the row's readers are done before the next row is staged
    }
    if (active) finish_particle<kXsph>(g, gi, st, L, lane, xv, vw);
}

}  // namespace

namespace pbfint {

pbf_status sweep_build_lists(const SweepGrid& sg, const pbf_float4* p, uint32_t* nbr, int* cnt,
                             size_t stride, uint32_t* status, void* stream) {
    if (sg.n == 0) return PBF_OK;
    if (stride < sg.n)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "neighbour list stride below particle count");
    if (sg.maxNeighbors < 0 || sg.maxNeighbors % 4 != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "grouped lists need a multiple of 4 slots");
    if (reinterpret_cast<uintptr_t>(nbr) % alignof(uint4) != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "grouped lists need 16-byte alignment");
    if (sg.n > 0xffffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "too many particles for 32-bit list entries");
    Grid g;
    g.p = p;
    g.cellStart = sg.cellStart;
    g.cellEnd = sg.cellEnd;
    g.sortedKeys = sg.sortedKeys;
    g.nbr = reinterpret_cast<uint4*>(nbr);
    g.cnt = cnt;
    g.status = status;
    g.n = sg.n;
    g.stride = stride;
    g.dims = sg.dims;
    g.h2 = sg.h * sg.h;
    g.maxNeighbors = sg.maxNeighbors;
    g.st = sg.st;
    g.skin = sg.skin;
    g.box = sg.box != nullptr ? *sg.box : pbf_container{};
    g.radius = sg.radius;
    g.fullBit = sg.skin ? PBF_STATUS_LIST_STALE : PBF_STATUS_NEIGHBOR_OVERFLOW;
    g.xs = sg.xs;
    g.vold = sg.vold;
    g.vout = sg.vout;
    g.tapeOut = sg.tapeOut;
    g.visc = sg.visc;
    g.inv_dt = sg.inv_dt;
    g.h = sg.h;
    if (sg.vout != nullptr && (sg.xs == nullptr || sg.vold == nullptr))
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "the XSPH build needs x and the velocities");
    if (sg.skin && sg.st.geom == nullptr && sg.box == nullptr)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "skin lists need the container");
    const size_t tasks = (sg.n + kThreads - 1) / kThreads;
    if (tasks > 0x7fffffffull)
        return pbf_fail(PBF_ERR_LENGTH, "too many particles for the neighbour sweep");
    const unsigned nb = static_cast<unsigned>(tasks);
    const bool batched = sg.st.geom != nullptr;
    if (sg.vout != nullptr) {
        if (sg.skin) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "XSPH is fused into exact builds only");
        if (batched) PBF_LAUNCH((k_build_lists<true, false, true>), nb, kThreads, stream, g);
        PBF_LAUNCH((k_build_lists<false, false, true>), nb, kThreads, stream, g);
    }
    if (batched && sg.skin) PBF_LAUNCH((k_build_lists<true, true, false>), nb, kThreads, stream, g);
    if (batched) PBF_LAUNCH((k_build_lists<true, false, false>), nb, kThreads, stream, g);
    if (sg.skin) PBF_LAUNCH((k_build_lists<false, true, false>), nb, kThreads, stream, g);
    PBF_LAUNCH((k_build_lists<false, false, false>), nb, kThreads, stream, g);
}

}  // namespace pbfint
