// Spatial grid for gfx950: cell keys, an in-house LDS-staged LSD radix sort
// of (key, slot) pairs, cellStart/cellEnd ranges and the neighbour search.
//
// Reference: src/pbf/spatial_grid.cu:100-226 (keys, CUB SortPairs, ranges),
// src/pbf/neighbors.cu:3-91 (findNeighbors).
//
// Radix sort design (per 8-bit digit pass, three launches):
//   1. k_radix_hist    one 1024-key tile per block; wave64 ballot-match
//                      aggregation, one LDS add per distinct digit per wave;
//                      writes counts[digit * tiles + tile].
//   2. k_digit_scan    one workgroup per digit scans that digit's counts
//                      across the tiles in place and writes its total.
//   3. k_radix_scatter scans the 256 digit totals in LDS (the digit's first
//                      output position), ranks keys stably inside the tile (rounds of 256
//                      keys in input order; rank = earlier rounds + earlier
//                      waves + earlier lanes with the same digit), stages
//                      the tile in LDS in digit order and writes it out so
//                      consecutive lanes store consecutive addresses.
// Only ceil(log2(numCells)) key bits are sorted: every key is below numCells,
// so the higher bits are zero and the order equals a 32-bit sort.
#include "pbf_device.hpp"
#include "pbf_internal.hpp"
#include "pbf_launch.hpp"

using namespace pbfdev;

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;  // 1024 keys per block
constexpr int kRadix = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kSortThreads * kScanItems;  // 4096 counters per block

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Lanes of this wave holding the same 8-bit digit (only lanes with valid set).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave,
                                                         uint32_t& total) {
    const unsigned lane = lane_id();
    const unsigned w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= static_cast<unsigned>(o)) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (unsigned q = 0; q < w; ++q) pre += s_wave[q];
    total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
    return pre + x - v;
}

__global__ void __launch_bounds__(kSortThreads)
    k_radix_hist(const uint32_t* __restrict__ keys, size_t n, int shift, uint32_t mask,
                 uint32_t* __restrict__ counts, uint32_t tiles) {
    __shared__ uint32_t s_hist[kRadix];
    const unsigned t = threadIdx.x;
    s_hist[t] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kSortTile;
    const uint64_t lt = (1ull << lane_id()) - 1ull;
#pragma unroll 4
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & mask : 0u;
        const uint64_t m = match_digit(d, valid);
        if (valid && (m & lt) == 0ull) atomicAdd(&s_hist[d], static_cast<uint32_t>(__popcll(m)));
    }
    __syncthreads();
    counts[static_cast<size_t>(t) * tiles + blockIdx.x] = s_hist[t];
}

__global__ void __launch_bounds__(kSortThreads)
    k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n, int shift,
                    uint32_t mask, const uint32_t* __restrict__ colPrefix, uint32_t tiles,
                    const uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_wcnt[4][kRadix];
    __shared__ uint32_t s_run[kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_base[kRadix];
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];

    const unsigned t = threadIdx.x;
    const unsigned w = t >> 6;
    const uint64_t lt = (1ull << lane_id()) - 1ull;
    const size_t base = static_cast<size_t>(blockIdx.x) * kSortTile;

    // Output position of each digit's first key: exclusive scan of the
    // digit totals (kRadix == kSortThreads, one digit per thread).
    {
        uint32_t all;
        s_base[t] = block_exclusive_scan(totals[t], s_wave, all);
    }
    s_run[t] = 0;
    s_wcnt[0][t] = 0;
    s_wcnt[1][t] = 0;
    s_wcnt[2][t] = 0;
    s_wcnt[3][t] = 0;

    uint32_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        const bool valid = idx < n;
        k[r] = valid ? kin[idx] : 0u;
        v[r] = valid ? (vin != nullptr ? vin[idx] : static_cast<uint32_t>(idx)) : 0u;
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        const bool valid = idx < n;
        const uint32_t d = (k[r] >> shift) & mask;
        const uint64_t m = match_digit(d, valid);
        const uint32_t pre = static_cast<uint32_t>(__popcll(m & lt));
        if (valid && pre == 0u) s_wcnt[w][d] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        if (valid) {
            uint32_t rk = s_run[d] + pre;
            for (unsigned q = 0; q < w; ++q) rk += s_wcnt[q][d];
            rank[r] = rk;
        }
        __syncthreads();
        s_run[t] += s_wcnt[0][t] + s_wcnt[1][t] + s_wcnt[2][t] + s_wcnt[3][t];
        s_wcnt[0][t] = 0;
        s_wcnt[1][t] = 0;
        s_wcnt[2][t] = 0;
        s_wcnt[3][t] = 0;
        __syncthreads();
    }

    uint32_t total;
    const uint32_t start = block_exclusive_scan(s_run[t], s_wave, total);
    s_start[t] = start;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        if (idx < n) {
            const uint32_t d = (k[r] >> shift) & mask;
            const uint32_t pos = s_start[d] + rank[r];
            s_key[pos] = k[r];
            s_val[pos] = v[r];
        }
    }
    __syncthreads();

    const size_t remaining = n - base;
    const uint32_t count = remaining < static_cast<size_t>(kSortTile)
                               ? static_cast<uint32_t>(remaining)
                               : static_cast<uint32_t>(kSortTile);
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t li = static_cast<uint32_t>(r) * kSortThreads + t;
        if (li < count) {
            const uint32_t kk = s_key[li];
            const uint32_t d = (kk >> shift) & mask;
            const size_t dst = static_cast<size_t>(s_base[d]) +
                               colPrefix[static_cast<size_t>(d) * tiles + blockIdx.x] +
                               (li - s_start[d]);
            kout[dst] = kk;
            vout[dst] = s_val[li];
        }
    }
}

// Padded LDS index for a 4096-element scan chunk (17 words per 16 values).
__device__ __forceinline__ unsigned scan_slot(unsigned i) { return (i >> 4) * 17u + (i & 15u); }

// Exclusive scan of a chunk of 4096 counters of src into dst (which may be
// src), plus a carry-in; with clearSrc, src's counters are zeroed once read.
__device__ void scan_chunk(uint32_t* src, uint32_t* dst, size_t base, size_t n, uint32_t carry,
                           bool clearSrc, uint32_t* s_buf, uint32_t* s_wave,
                           uint32_t& chunkTotal) {
    const unsigned t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const unsigned li = static_cast<unsigned>(r) * kSortThreads + t;
        const size_t idx = base + li;
        s_buf[scan_slot(li)] = idx < n ? src[idx] : 0u;
        if (clearSrc && idx < n) src[idx] = 0u;
    }
    __syncthreads();
    uint32_t vals[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        vals[q] = s_buf[scan_slot(t * kScanItems + q)];
        sum += vals[q];
    }
    uint32_t total;
    uint32_t run = carry + block_exclusive_scan(sum, s_wave, total);
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        s_buf[scan_slot(t * kScanItems + q)] = run;
        run += vals[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const unsigned li = static_cast<unsigned>(r) * kSortThreads + t;
        const size_t idx = base + li;
        if (idx < n) dst[idx] = s_buf[scan_slot(li)];
    }
    __syncthreads();
    chunkTotal = total;
}

// Exclusive scan of one digit's counts across the tiles, in place (one
// workgroup per digit), and the digit's total.
__global__ void __launch_bounds__(kSortThreads)
    k_digit_scan(uint32_t* __restrict__ counts, uint32_t tiles, uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_buf[kScanChunk / 16 * 17];
    __shared__ uint32_t s_wave[4];
    uint32_t* col = counts + static_cast<size_t>(blockIdx.x) * tiles;
    uint32_t carry = 0;
    for (size_t b = 0; b < tiles; b += kScanChunk) {
        uint32_t total;
        scan_chunk(col, col, b, tiles, carry, false, s_buf, s_wave, total);
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// ---------------------------------------------------------------------------
// Re-sort after a small move.  The keys of a build are those of particles
// that sit in the order of the previous build's sort, so prev[i] (that
// sort's output) is slot i's previous key and prev is non-decreasing.  The
// slots whose key did not change ("kept") are therefore already in (key,
// slot) order; the changed ones are sorted on their own and merged in.  The
// result is the stable sort by key for any non-decreasing prev; only the
// cost depends on how many keys changed (about 4% per constraint
// iteration on the dam breaks, 20% after the prediction).
//
// The merge is split by tiles of kResortTile slots of the previous order.
// Splitter t is (prev[t*T], t*T) (splitters[t] holds prev[t*T]); the kept
// slots of tile t lie between splitters t and t+1, so each changed entry
// joins the segment of the tile whose splitter interval holds it, and the
// output of tile t is the merge of its kept slots with its segment,
// written from (t*T - changed slots before tile t) + (entries of the
// segments before t).
//   k_resort_count    marks the changed slots (bit mask), finds their
//                     segments (kept in the output index array until the
//                     merge overwrites it) and counts them per tile and per
//                     segment;
//   k_resort_scan     one workgroup: exclusive scans of both counts, clears
//                     the counters for the next re-sort;
//   k_resort_scatter  writes each changed (key, slot) into its segment (in
//                     no particular order inside it);
//   k_resort_merge    per tile: sorts its segment by (key, slot) (ranks by
//                     counting up to 256 entries, bitonic in LDS up to
//                     kResortCap, bitonic in place in global memory beyond),
//                     merges it with the kept slots by binary searches and
//                     writes both into the tile's output range.
// The two counting kernels first group a workgroup's changed entries by
// segment in an LDS table, so that each (workgroup, segment) pair costs one
// device-scope atomic: such atomics on counters that many workgroups share
// serialise, and one per entry made the count pass five times slower.
constexpr uint32_t kResortTile = kSortTile;
constexpr uint32_t kResortCap = 1024;
constexpr uint32_t kGroupSlots = kSortThreads;  // distinct segments a workgroup groups in LDS
constexpr uint32_t kNoSegment = 0xFFFFFFFFu;
constexpr uint32_t kCoarseMax = 1024;  // splitters of the coarse search index in LDS
constexpr int kResortScanThreads = 1024;

// The re-sort's counters in the grid's scratch, laid out for the grid's
// capacity (not the build's count) so that the zeroed counters of a count
// never overlap words another count's radix sort or re-sort has written.
// Every array starts on a 16-byte boundary.
struct ResortScratch {
    uint32_t* segCount;   // tiles: changed entries per segment (zero between re-sorts)
    uint32_t* segFill;    // tiles: entries placed so far per segment
    uint32_t* segStart;   // tiles + 1: exclusive scan of segCount
    uint32_t* chgCount;   // tiles: changed slots per tile
    uint32_t* chgBefore;  // tiles + 1: exclusive scan of chgCount
    uint64_t* mask;       // ceil(n / 64): bit i of word i/64 set for a changed slot
};

__host__ __device__ inline size_t resort_tiles(size_t n) {
    return (n + kResortTile - 1) / kResortTile;
}

// Words of one counter array: tiles + 1, rounded up to a multiple of 4.
__host__ __device__ inline size_t resort_stride(size_t n) { return (resort_tiles(n) + 4) & ~size_t{3}; }

__host__ __device__ inline size_t resort_words(size_t n) {
    return 5 * resort_stride(n) + 2 * ((n + 63) / 64);
}

inline ResortScratch resort_layout(uint32_t* w, size_t capacity) {
    const size_t S = resort_stride(capacity);
    ResortScratch r;
    r.segCount = w;
    r.segFill = w + S;
    r.segStart = w + 2 * S;
    r.chgCount = w + 3 * S;
    r.chgBefore = w + 4 * S;
    r.mask = reinterpret_cast<uint64_t*>(w + 5 * S);
    return r;
}

__device__ __forceinline__ uint64_t key_slot(uint32_t key, size_t slot) {
    return (static_cast<uint64_t>(key) << 32) | static_cast<uint64_t>(slot);
}

// Whether splitter t, (sk, t*T), is not above (key, slot).
__device__ __forceinline__ bool splitter_not_above(uint32_t sk, uint32_t t, uint32_t key,
                                                   size_t slot) {
    return sk < key || (sk == key && static_cast<size_t>(t) * kResortTile <= slot);
}

// A workgroup's changed entries grouped by segment: an open-addressing
// table in LDS (a slot, once claimed for a segment, keeps it).
struct SegmentGroups {
    uint32_t seg[kGroupSlots];
    uint32_t count[kGroupSlots];
    uint32_t base[kGroupSlots];
};

__device__ __forceinline__ void groups_clear(SegmentGroups& g) {
    g.seg[threadIdx.x] = kNoSegment;
    g.count[threadIdx.x] = 0u;
}

// Adds one entry of segment s: its group and its rank in the group, or -1
// when the table has no slot left for s.
__device__ __forceinline__ int groups_add(SegmentGroups& g, uint32_t s, uint32_t& rank) {
    const uint32_t h = (s * 0x9E3779B1u) >> 24;
    for (uint32_t probe = 0; probe < kGroupSlots; ++probe) {
        const uint32_t q = (h + probe) & (kGroupSlots - 1);
        const uint32_t cur = atomicCAS(&g.seg[q], kNoSegment, s);
        if (cur == kNoSegment || cur == s) {
            rank = atomicAdd(&g.count[q], 1u);
            return static_cast<int>(q);
        }
    }
    return -1;
}

// The segment of each changed entry is the last tile t whose splitter is
// not above it (tile 0 has no lower bound): a binary search over every
// stride-th splitter in LDS, then over the stride splitters it leaves, in
// global memory, the workgroup's four entries per thread side by side.
__global__ void __launch_bounds__(kSortThreads)
    k_resort_count(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ prev,
                   const uint32_t* __restrict__ splitters, size_t n, uint32_t tiles,
                   ResortScratch rs, uint32_t* __restrict__ segOf) {
    __shared__ SegmentGroups g;
    __shared__ uint32_t s_coarse[kCoarseMax];
    __shared__ uint32_t s_changed;
    const unsigned t = threadIdx.x;
    const uint32_t stride =
        tiles <= 16u * kCoarseMax ? 16u : (tiles + kCoarseMax - 1) / kCoarseMax;
    const uint32_t coarse = (tiles + stride - 1) / stride;
    for (uint32_t c = t; c < coarse; c += kSortThreads) s_coarse[c] = splitters[c * stride];
    groups_clear(g);
    if (t == 0) s_changed = 0;

    const size_t base = static_cast<size_t>(blockIdx.x) * kResortTile;
    uint32_t k[kSortItems];
    bool ch[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        const bool valid = idx < n;
        k[r] = valid ? keys[idx] : 0u;
        ch[r] = valid && k[r] != prev[idx];
    }
    __syncthreads();

    uint32_t mine = 0;
    uint32_t lo[kSortItems], hi[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        const uint64_t b = __ballot(ch[r]);
        // Lane 0 sits at a multiple of 64 slots: its word of the mask.
        if (lane_id() == 0 && idx < n) rs.mask[idx >> 6] = b;
        mine += static_cast<uint32_t>(__popcll(b));
        lo[r] = hi[r] = 0u;
        if (!ch[r]) continue;
        uint32_t a = 0, z = coarse - 1;
        while (a < z) {
            const uint32_t mid = (a + z + 1) >> 1;
            if (splitter_not_above(s_coarse[mid], mid * stride, k[r], idx))
                a = mid;
            else
                z = mid - 1;
        }
        lo[r] = a * stride;
        hi[r] = (lo[r] + stride < tiles ? lo[r] + stride : tiles) - 1u;
    }
    for (;;) {
        bool open = false;
        uint32_t sk[kSortItems];
#pragma unroll
        for (int r = 0; r < kSortItems; ++r)
            sk[r] = lo[r] < hi[r] ? splitters[(lo[r] + hi[r] + 1) >> 1] : 0u;
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            if (lo[r] >= hi[r]) continue;
            const uint32_t mid = (lo[r] + hi[r] + 1) >> 1;
            const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
            if (splitter_not_above(sk[r], mid, k[r], idx))
                lo[r] = mid;
            else
                hi[r] = mid - 1;
            open = open || lo[r] < hi[r];
        }
        if (!open) break;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (!ch[r]) continue;
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        segOf[idx] = lo[r];
        uint32_t rank;
        if (groups_add(g, lo[r], rank) < 0) atomicAdd(&rs.segCount[lo[r]], 1u);
    }
    if (lane_id() == 0 && mine != 0u) atomicAdd(&s_changed, mine);
    __syncthreads();
    if (g.seg[t] != kNoSegment) atomicAdd(&rs.segCount[g.seg[t]], g.count[t]);
    if (t == 0) rs.chgCount[blockIdx.x] = s_changed;
}

// Exclusive scans of segCount (into segStart) and chgCount (into
// chgBefore), clearing segCount and segFill for the next re-sort: chunks
// of 4096 tiles, four consecutive tiles of each array per thread (16-byte
// loads and stores; the arrays are padded to a multiple of four words).
__global__ void __launch_bounds__(kResortScanThreads) k_resort_scan(uint32_t tiles,
                                                                    ResortScratch rs) {
    constexpr unsigned kWaves = kResortScanThreads / 64;
    __shared__ uint32_t s_wave[2][kWaves];
    const unsigned t = threadIdx.x;
    const unsigned lane = lane_id();
    const unsigned w = t >> 6;
    uint32_t carryA = 0, carryC = 0;
    for (uint32_t chunk = 0; chunk < tiles; chunk += 4 * kResortScanThreads) {
        const uint32_t i0 = chunk + 4 * t;
        uint32_t a[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
        if (i0 < tiles) {
            const uint4 va = *reinterpret_cast<const uint4*>(rs.segCount + i0);
            const uint4 vc = *reinterpret_cast<const uint4*>(rs.chgCount + i0);
            a[0] = va.x; a[1] = va.y; a[2] = va.z; a[3] = va.w;
            c[0] = vc.x; c[1] = vc.y; c[2] = vc.z; c[3] = vc.w;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (i0 + q >= tiles) a[q] = c[q] = 0u;  // words past the tiles
        }
        const uint32_t sa = a[0] + a[1] + a[2] + a[3], sc = c[0] + c[1] + c[2] + c[3];
        uint32_t ia = sa, ic = sc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t ya = __shfl_up(ia, o, 64), yc = __shfl_up(ic, o, 64);
            if (lane >= static_cast<unsigned>(o)) {
                ia += ya;
                ic += yc;
            }
        }
        if (lane == 63) {
            s_wave[0][w] = ia;
            s_wave[1][w] = ic;
        }
        __syncthreads();
        uint32_t pa = carryA + ia - sa, pc = carryC + ic - sc, ta = 0, tc = 0;
        for (unsigned q = 0; q < kWaves; ++q) {
            if (q < w) {
                pa += s_wave[0][q];
                pc += s_wave[1][q];
            }
            ta += s_wave[0][q];
            tc += s_wave[1][q];
        }
        __syncthreads();
        if (i0 < tiles) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (i0 + q >= tiles) break;
                rs.segStart[i0 + q] = pa;
                rs.chgBefore[i0 + q] = pc;
                rs.segCount[i0 + q] = 0u;
                rs.segFill[i0 + q] = 0u;
                pa += a[q];
                pc += c[q];
            }
        }
        carryA += ta;
        carryC += tc;
    }
    if (t == 0) {
        rs.segStart[tiles] = carryA;
        rs.chgBefore[tiles] = carryC;
    }
}

__global__ void __launch_bounds__(kSortThreads)
    k_resort_scatter(const uint32_t* __restrict__ keys, size_t n, ResortScratch rs,
                     const uint32_t* __restrict__ segOf, uint32_t* __restrict__ segKey,
                     uint32_t* __restrict__ segSlot) {
    __shared__ SegmentGroups g;
    const unsigned t = threadIdx.x;
    groups_clear(g);
    const size_t base = static_cast<size_t>(blockIdx.x) * kResortTile;
    uint32_t key[kSortItems], seg[kSortItems], rank[kSortItems];
    bool ch[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        ch[r] = idx < n && ((rs.mask[idx >> 6] >> (idx & 63)) & 1ull) != 0ull;
        key[r] = ch[r] ? keys[idx] : 0u;
        seg[r] = ch[r] ? segOf[idx] : 0u;
    }
    __syncthreads();
    int grp[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        grp[r] = -2;  // not a changed slot
        rank[r] = 0u;
        if (!ch[r]) continue;
        grp[r] = groups_add(g, seg[r], rank[r]);
        if (grp[r] == -1) {
            const uint32_t pos = rs.segStart[seg[r]] + atomicAdd(&rs.segFill[seg[r]], 1u);
            segKey[pos] = key[r];
            segSlot[pos] = static_cast<uint32_t>(base + static_cast<size_t>(r) * kSortThreads + t);
        }
    }
    __syncthreads();
    if (g.seg[t] != kNoSegment)
        g.base[t] = rs.segStart[g.seg[t]] + atomicAdd(&rs.segFill[g.seg[t]], g.count[t]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (grp[r] < 0) continue;
        const uint32_t pos = g.base[grp[r]] + rank[r];
        segKey[pos] = key[r];
        segSlot[pos] = static_cast<uint32_t>(base + static_cast<size_t>(r) * kSortThreads + t);
    }
}

// Ascending bitonic sort of s values, by the workgroup, on a power-of-two
// network of P >= s inputs whose inputs from s on are taken as +infinity:
// every comparator puts the smaller value at the lower index, so those
// inputs never move and their comparators are skipped.  get(i) reads
// value i, put(i, v) writes it; a barrier separates the stages.
template <class Get, class Put>
__device__ void bitonic_sort(uint32_t s, Get get, Put put) {
    if (s < 2) return;
    uint32_t P = 2;
    while (P < s) P <<= 1;
    const uint32_t pairs = P >> 1;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t d = k >> 1; d > 0; d >>= 1) {
            for (uint32_t q = threadIdx.x; q < pairs; q += blockDim.x) {
                uint32_t i, j;
                if (d == (k >> 1)) {  // mirror comparators across the block of k
                    const uint32_t blk = q / d, o = q % d;
                    i = blk * k + o;
                    j = blk * k + k - 1 - o;
                } else {  // half-cleaners at distance d
                    i = (q / d) * 2 * d + q % d;
                    j = i + d;
                }
                if (j < s) {
                    const uint64_t a = get(i), b = get(j);
                    if (b < a) {
                        put(i, b);
                        put(j, a);
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Entries of v[0, len) below x (v ascending).
template <class Get>
__device__ __forceinline__ uint32_t count_below(Get v, uint32_t len, uint64_t x) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v(mid) < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kSortThreads)
    k_resort_merge(const uint32_t* __restrict__ keys, size_t n, ResortScratch rs,
                   uint32_t* __restrict__ segKey, uint32_t* __restrict__ segSlot,
                   uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ uint64_t s_kept[kResortTile];
    __shared__ uint64_t s_seg[kResortCap];
    __shared__ uint32_t s_cnt[kSortItems][4];

    const unsigned t = threadIdx.x;
    const unsigned w = t >> 6;
    const uint64_t lt = (1ull << lane_id()) - 1ull;
    const uint32_t tile = blockIdx.x;
    const size_t base = static_cast<size_t>(tile) * kResortTile;
    const uint32_t s0 = rs.segStart[tile];
    const uint32_t s = rs.segStart[tile + 1] - s0;
    const size_t out0 = base - rs.chgBefore[tile] + s0;
    const bool inLds = s <= kResortCap;

    // Kept slots of the tile, in slot order, into s_kept.
    uint64_t kv[kSortItems];
    uint32_t rank[kSortItems];
    bool kept[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const size_t idx = base + static_cast<size_t>(r) * kSortThreads + t;
        kept[r] = idx < n && ((rs.mask[idx >> 6] >> (idx & 63)) & 1ull) == 0ull;
        kv[r] = kept[r] ? key_slot(keys[idx], idx) : 0ull;
        const uint64_t b = __ballot(kept[r]);
        rank[r] = static_cast<uint32_t>(__popcll(b & lt));
        if (lane_id() == 0) s_cnt[r][w] = static_cast<uint32_t>(__popcll(b));
    }
    if (inLds)
        for (uint32_t j = t; j < s; j += kSortThreads)
            s_seg[j] = (static_cast<uint64_t>(segKey[s0 + j]) << 32) | segSlot[s0 + j];
    __syncthreads();
    uint32_t nKept = 0;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        for (unsigned q = 0; q < 4; ++q) {
            if (q == w) rank[r] += nKept;
            nKept += s_cnt[r][q];
        }
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
        if (kept[r]) s_kept[rank[r]] = kv[r];

    // The segment in (key, slot) order (its entries are distinct).
    uint32_t* gk = segKey + s0;
    uint32_t* gs = segSlot + s0;
    const auto gseg = [&](uint32_t i) { return (static_cast<uint64_t>(gk[i]) << 32) | gs[i]; };
    if (inLds && s <= kSortThreads) {
        uint64_t v = 0ull;
        uint32_t rk = 0u;
        if (t < s) {
            v = s_seg[t];
            for (uint32_t i = 0; i < s; ++i) rk += s_seg[i] < v ? 1u : 0u;
        }
        __syncthreads();
        if (t < s) s_seg[rk] = v;
    } else if (inLds) {
        bitonic_sort(
            s, [&](uint32_t i) { return s_seg[i]; }, [&](uint32_t i, uint64_t v) { s_seg[i] = v; });
    } else {
        bitonic_sort(s, gseg, [&](uint32_t i, uint64_t v) {
            gk[i] = static_cast<uint32_t>(v >> 32);
            gs[i] = static_cast<uint32_t>(v);
        });
    }
    __syncthreads();  // s_kept and the sorted segment complete

    // Output position = the tile's start + entries of the other list below.
    const auto lseg = [&](uint32_t i) { return s_seg[i]; };
    const auto lkept = [&](uint32_t i) { return s_kept[i]; };
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (!kept[r]) continue;
        const uint32_t below = inLds ? count_below(lseg, s, kv[r]) : count_below(gseg, s, kv[r]);
        const size_t o = out0 + rank[r] + below;
        kout[o] = static_cast<uint32_t>(kv[r] >> 32);
        vout[o] = static_cast<uint32_t>(kv[r]);
    }
    for (uint32_t j = t; j < s; j += kSortThreads) {
        const uint64_t v = inLds ? s_seg[j] : gseg(j);
        const size_t o = out0 + j + count_below(lkept, nKept, v);
        kout[o] = static_cast<uint32_t>(v >> 32);
        vout[o] = static_cast<uint32_t>(v);
    }
}

// Keys with a single cell: the sort is the identity.
__global__ void __launch_bounds__(kBlock) k_sort_identity(const uint32_t* __restrict__ kin,
                                                          uint32_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kout[i] = kin[i];
    vout[i] = static_cast<uint32_t>(i);
}

// cudaMemset(-1) of cellStart and cellEnd (spatial_grid.cu:203-204); ce may
// be null when only cellStart has readers of empty cells.
__global__ void __launch_bounds__(kBlock) k_fill_cells(int* __restrict__ cs, int* __restrict__ ce,
                                                       size_t cells) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < cells;
         i += stride) {
        cs[i] = -1;
        if (ce != nullptr) ce[i] = -1;
    }
}

// buildCellRanges (spatial_grid.cu:123-137): half-open [start, end).  With
// splitters not null, also the keys at multiples of kResortTile, which the
// next re-sort splits its merge by.
__global__ void __launch_bounds__(kBlock) k_cell_ranges(const uint32_t* __restrict__ sk,
                                                        int* __restrict__ cs, int* __restrict__ ce,
                                                        size_t n, uint32_t* __restrict__ splitters) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sk[i];
    if (splitters != nullptr && i % kResortTile == 0) splitters[i / kResortTile] = k;
    if (i == 0 || sk[i - 1] != k) cs[k] = static_cast<int>(i);
    if (i == n - 1 || sk[i + 1] != k) ce[k] = static_cast<int>(i + 1);
}

// The cell ranges of the sorted keys (k_cell_ranges) and, in the same
// pass, the permutation of the working set into that order
// (gatherSolverState, src/pbf/pbf_solver_kernels.cu:387-407): slot i of the
// outputs takes the records of slot idx[i].  Workgroups take their slots
// contiguously per XCD, as the passes around it do: a slot's record mostly
// comes from a nearby slot that the same XCD wrote or reads next (1M dam
// break: 15.7 against 17.5 us per launch with the plain order).
__global__ void __launch_bounds__(kBlock)
    k_ranges_gather(const uint32_t* __restrict__ sk, int* __restrict__ cs, int* __restrict__ ce,
                    size_t n, pbfint::Permute pm, uint32_t* __restrict__ splitters) {
    const size_t i =
        static_cast<size_t>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = pm.idx[i];
    st4(pm.po, i, ld4(pm.p, s));
    const uint32_t r = pm.permIn != nullptr ? pm.permIn[s] : s;
    if (pm.x != nullptr) {
        st4(pm.xo, i, ld4(pm.x, r));
        st4(pm.vo, i, ld4(pm.v, r));
        st4(pm.vino, i, ld4(pm.vin, r));
        pm.ido[i] = pm.id[r];
    } else {
        pm.permOut[i] = r;
    }
    const uint32_t k = sk[i];
    if (splitters != nullptr && i % kResortTile == 0) splitters[i / kResortTile] = k;
    if (i == 0 || sk[i - 1] != k) cs[k] = static_cast<int>(i);
    if (i == n - 1 || sk[i + 1] != k) ce[k] = static_cast<int>(i + 1);
}

// findNeighbors (src/pbf/neighbors.cu:3-91).  Cells are swept z, y, x over
// [cell - R, cell + R] clamped to the grid, particles of a cell in sorted
// order; the radius test is inclusive and self is excluded by index.
__global__ void __launch_bounds__(kBlock)
    k_find_neighbors(const pbf_float4* __restrict__ p, const int* __restrict__ cs,
                     const int* __restrict__ ce, pbf_int3 g, pbf_float3 mn, float cellSize,
                     size_t n, float h, long long R, uint32_t* __restrict__ nbr,
                     int* __restrict__ cnt, int maxN, size_t stride,
                     uint32_t* __restrict__ flag, uint32_t bit,
                     const uint32_t* __restrict__ guard) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (guard != nullptr && *guard != 0u) return;
    const float4 pi = ld4(p, i);
    const int3 c = cell_of(pi, mn, cellSize, g);
    const float h2 = h * h;
    const long long zlo = c.z - R < 0 ? 0 : c.z - R;
    const long long ylo = c.y - R < 0 ? 0 : c.y - R;
    const long long xlo = c.x - R < 0 ? 0 : c.x - R;
    const long long zhi = c.z + R > g.z - 1 ? g.z - 1 : c.z + R;
    const long long yhi = c.y + R > g.y - 1 ? g.y - 1 : c.y + R;
    const long long xhi = c.x + R > g.x - 1 ? g.x - 1 : c.x + R;
    int count = 0;
    for (long long z = zlo; z <= zhi; ++z) {
        for (long long y = ylo; y <= yhi; ++y) {
            for (long long x = xlo; x <= xhi; ++x) {
                const uint32_t key = key_of(static_cast<int>(x), static_cast<int>(y),
                                            static_cast<int>(z), g);
                const int s = cs[key];
                if (s == -1) continue;
                const int e = ce[key];
                for (int j = s; j < e; ++j) {
                    if (static_cast<size_t>(j) == i) continue;
                    const float4 pj = ld4(p, j);
                    const float d2 = dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                    if (d2 <= h2) {
                        if (count < maxN) nbr[static_cast<size_t>(count) * stride + i] = static_cast<uint32_t>(j);
                        ++count;
                    }
                }
            }
        }
    }
    cnt[i] = count;
    if (count > maxN && flag != nullptr) atomicOr(flag, bit);
}

}  // namespace

namespace pbfint {

// Stencil radius ceil(h / cellSize) evaluated in double like the reference
// (neighbors.cu:39), saturated so the clamp to the grid stays exact.
long long stencil_radius(float h, float cellSize) {
    const double r = __builtin_ceil(static_cast<double>(h) / static_cast<double>(cellSize));
    if (!(r >= 0.0)) return 0;
    if (r > 4.0e18) return 4000000000000000000LL;
    return static_cast<long long>(r);
}

pbf_status launch_find_neighbors(const pbf_float4* p, const int* cs, const int* ce, pbf_int3 g,
                                 pbf_float3 mn, float cellSize, size_t n, float h,
                                 uint32_t* nbr, int* cnt, int maxN, size_t stride,
                                 uint32_t* flag, uint32_t bit, const uint32_t* guard,
                                 void* stream) {
    if (n == 0) return PBF_OK;
    const long long R = stencil_radius(h, cellSize);
    PBF_LAUNCH(k_find_neighbors, grid_for(n), kBlock, stream, p, cs, ce, g, mn, cellSize, n, h, R,
               nbr, cnt, maxN, stride, flag, bit, guard);
}

pbf_status launch_fill_cells(int* cs, int* ce, size_t cells, void* stream) {
    if (cells == 0) return PBF_OK;
    size_t blocks = (cells + kBlock - 1) / kBlock;
    if (blocks > 16384) blocks = 16384;
    PBF_LAUNCH(k_fill_cells, static_cast<unsigned>(blocks), kBlock, stream, cs, ce, cells);
}

pbf_status launch_cell_ranges(const uint32_t* sk, int* cs, int* ce, size_t n, void* stream,
                              uint32_t* splitters) {
    if (n == 0) return PBF_OK;
    PBF_LAUNCH(k_cell_ranges, grid_for(n), kBlock, stream, sk, cs, ce, n, splitters);
}

pbf_status launch_ranges_gather(const uint32_t* sk, int* cs, int* ce, size_t n, const Permute& pm,
                                void* stream, uint32_t* splitters) {
    if (n == 0) return PBF_OK;
    if (pm.x == nullptr && pm.permOut == nullptr)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "a p*-only permutation needs permOut");
    if (pm.permOut != nullptr && pm.permOut == pm.permIn)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "permutation composed in place");
    PBF_LAUNCH(k_ranges_gather, grid_for(n), kBlock, stream, sk, cs, ce, n, pm, splitters);
}

size_t radix_scratch_words(size_t n) {
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    return tiles * kRadix + kRadix + 64;
}

// The radix sort's words, then (from a 256-byte boundary) the re-sort's, for
// a capacity of n keys.
static size_t resort_offset(size_t n) { return (radix_scratch_words(n) + 63) & ~size_t{63}; }

size_t grid_scratch_words(size_t n) { return resort_offset(n) + resort_words(n); }

size_t resort_splitter_count(size_t n) { return resort_tiles(n); }

// The re-sort of keys given prev, the previous sort's keys of the same
// slots (non-decreasing), and its splitters (prev at multiples of the tile;
// launch_cell_ranges and launch_ranges_gather write them): kout/vout
// receive the stable sort of (keys, slot) exactly as radix_sort_pairs
// gives it.  kout may be prev (prev is read before kout is written);
// vout holds the changed slots' segments until the merge writes it;
// tmpK/tmpV hold the changed entries.  scratch is the grid's
// (grid_scratch_words(capacity), capacity >= n), whose re-sort counters are
// zero between calls.
pbf_status resort_pairs(const uint32_t* keys, const uint32_t* prev, const uint32_t* splitters,
                        uint32_t* kout, uint32_t* vout, uint32_t* tmpK, uint32_t* tmpV, size_t n,
                        size_t capacity, uint32_t* scratch, void* stream) {
    if (n == 0) return PBF_OK;
    if (capacity < n) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "re-sort beyond capacity");
    if (keys == kout || keys == tmpK || keys == tmpV || prev == tmpK || prev == tmpV ||
        kout == tmpK || kout == tmpV || vout == tmpK || vout == tmpV || tmpK == tmpV ||
        kout == vout || keys == vout || prev == vout)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "re-sort buffers must not alias");
    if (resort_tiles(n) > 0xFFFFFFFFull)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "re-sort tile count exceeds 32 bits");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t tiles = static_cast<uint32_t>(resort_tiles(n));
    const ResortScratch rs = resort_layout(scratch + resort_offset(capacity), capacity);
    hipLaunchKernelGGL(k_resort_count, dim3(tiles), dim3(kSortThreads), 0, s, keys, prev,
                       splitters, n, tiles, rs, vout);
    hipLaunchKernelGGL(k_resort_scan, dim3(1), dim3(kResortScanThreads), 0, s, tiles, rs);
    hipLaunchKernelGGL(k_resort_scatter, dim3(tiles), dim3(kSortThreads), 0, s, keys, n, rs, vout,
                       tmpK, tmpV);
    hipLaunchKernelGGL(k_resort_merge, dim3(tiles), dim3(kSortThreads), 0, s, keys, n, rs, tmpK,
                       tmpV, kout, vout);
    return pbf_hip_check(hipGetLastError());
}

int radix_key_bits(size_t numCells) {
    int bits = 0;
    while (bits < 32 && (static_cast<uint64_t>(1) << bits) < static_cast<uint64_t>(numCells)) ++bits;
    return bits;
}

// Stable sort of (key, slot) pairs by the low `bits` key bits.  vin ==
// nullptr means the values are the slot indices 0..n-1.  The passes run
// from the inputs through tmpK/tmpV and kout/vout in turn, ending in
// kout/vout; a pass never writes the array it reads (its tiles are read
// and written by different blocks), so the five arrays must be distinct.
// scratch holds radix_scratch_words(n) words.
pbf_status radix_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                            uint32_t* vout, uint32_t* tmpK, uint32_t* tmpV, size_t n, int bits,
                            uint32_t* scratch, void* stream) {
    if (n == 0) return PBF_OK;
    if (kin == tmpK || kin == kout || (vin != nullptr && (vin == tmpV || vin == vout)) ||
        kout == tmpK || vout == tmpV)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "radix sort buffers must not alias");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (bits <= 0) {
        hipLaunchKernelGGL(k_sort_identity, dim3(grid_for(n)), dim3(kBlock), 0, s, kin, kout, vout,
                           n);
        return pbf_hip_check(hipGetLastError());
    }
    const int passes = (bits + 7) / 8;
    const uint32_t tiles = static_cast<uint32_t>((n + kSortTile - 1) / kSortTile);
    const size_t counters = static_cast<size_t>(tiles) * kRadix;
    uint32_t* counts = scratch;
    uint32_t* totals = scratch + counters;

    const uint32_t* srcK = kin;
    const uint32_t* srcV = vin;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = pass * 8;
        const int width = bits - shift < 8 ? bits - shift : 8;
        const uint32_t mask = (1u << width) - 1u;
        const bool toOut = ((passes - 1 - pass) % 2) == 0;
        uint32_t* dstK = toOut ? kout : tmpK;
        uint32_t* dstV = toOut ? vout : tmpV;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortThreads), 0, s, srcK, n, shift,
                           mask, counts, tiles);
        hipLaunchKernelGGL(k_digit_scan, dim3(kRadix), dim3(kSortThreads), 0, s, counts, tiles,
                           totals);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, srcK, srcV,
                           dstK, dstV, n, shift, mask, counts, tiles, totals);
        const pbf_status st = pbf_hip_check(hipGetLastError());
        if (st != PBF_OK) return st;
        srcK = dstK;
        srcV = dstV;
    }
    return PBF_OK;
}

}  // namespace pbfint

extern "C" pbf_status pbf_k_find_neighbors(const pbf_float4* p, const int* cs, const int* ce,
                                           pbf_int3 g, pbf_float3 mn, float cellSize, size_t n,
                                           float h, uint32_t* nbr, int* cnt, int maxN,
                                           size_t stride, int* flag, void* stream) {
    return pbfint::launch_find_neighbors(p, cs, ce, g, mn, cellSize, n, h, nbr, cnt, maxN, stride,
                                         reinterpret_cast<uint32_t*>(flag), 1u, nullptr, stream);
}
