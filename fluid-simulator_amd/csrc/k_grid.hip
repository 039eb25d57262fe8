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

// Exclusive scan of a chunk of 4096 counters in place, plus a carry-in.
__device__ void scan_chunk(uint32_t* __restrict__ data, size_t base, size_t n, uint32_t carry,
                           uint32_t* s_buf, uint32_t* s_wave, uint32_t& chunkTotal) {
    const unsigned t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        const unsigned li = static_cast<unsigned>(r) * kSortThreads + t;
        const size_t idx = base + li;
        s_buf[scan_slot(li)] = idx < n ? data[idx] : 0u;
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
        if (idx < n) data[idx] = s_buf[scan_slot(li)];
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
        scan_chunk(col, b, tiles, carry, s_buf, s_wave, total);
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
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

// Neighbour-sweep tile (kSweepTileW cells of one grid row) of key k at
// slot i: in a batched step, the tile of the slot's scene, offset into the
// global tile range.
template <bool kBatched>
__device__ __forceinline__ uint32_t sweep_tile_of(uint32_t k, size_t i, uint32_t gx,
                                                  uint32_t chunks, const pbfint::SceneTable& st) {
    if (kBatched) {
        const pbfint::SceneGeom& sg = st.geom[st.sceneOf[i]];
        const uint32_t lk = k - sg.cellOff;
        const uint32_t sgx = static_cast<uint32_t>(sg.dims.x);
        return sg.tileOff + (lk / sgx) * sg.chunks + (lk % sgx) / pbfint::kSweepTileW;
    }
    return (k / gx) * chunks + (k % gx) / pbfint::kSweepTileW;
}

// buildCellRanges (spatial_grid.cu:123-137): half-open [start, end).
// With tileEnd not null, also records one past the last particle of every
// non-empty neighbour-sweep tile, which the last particle of the tile
// writes.
template <bool kBatched>
__global__ void __launch_bounds__(kBlock) k_cell_ranges(const uint32_t* __restrict__ sk,
                                                        int* __restrict__ cs, int* __restrict__ ce,
                                                        size_t n, uint32_t gx, uint32_t chunks,
                                                        uint32_t* __restrict__ tileEnd,
                                                        pbfint::SceneTable st) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = sk[i];
    if (i == 0 || sk[i - 1] != k) cs[k] = static_cast<int>(i);
    const bool lastInCell = i == n - 1 || sk[i + 1] != k;
    if (!lastInCell) return;
    ce[k] = static_cast<int>(i + 1);
    if (tileEnd != nullptr) {
        const uint32_t t = sweep_tile_of<kBatched>(k, i, gx, chunks, st);
        if (i == n - 1 || sweep_tile_of<kBatched>(sk[i + 1], i + 1, gx, chunks, st) != t)
            tileEnd[t] = static_cast<uint32_t>(i + 1);
    }
}

// The cell ranges of the sorted keys (k_cell_ranges) and, in the same
// pass, the permutation of the working set into that order
// (gatherSolverState, src/pbf/pbf_solver_kernels.cu:387-407): slot i of the
// outputs takes the records of slot idx[i].
template <bool kBatched>
__global__ void __launch_bounds__(kBlock)
    k_ranges_gather(const uint32_t* __restrict__ sk, int* __restrict__ cs, int* __restrict__ ce,
                    size_t n, uint32_t gx, uint32_t chunks, uint32_t* __restrict__ tileEnd,
                    pbfint::SceneTable st, pbfint::Permute pm) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
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
    if (i == 0 || sk[i - 1] != k) cs[k] = static_cast<int>(i);
    const bool lastInCell = i == n - 1 || sk[i + 1] != k;
    if (!lastInCell) return;
    ce[k] = static_cast<int>(i + 1);
    const uint32_t t = sweep_tile_of<kBatched>(k, i, gx, chunks, st);
    if (i == n - 1 || sweep_tile_of<kBatched>(sk[i + 1], i + 1, gx, chunks, st) != t)
        tileEnd[t] = static_cast<uint32_t>(i + 1);
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

pbf_status launch_cell_ranges(const uint32_t* sk, int* cs, int* ce, size_t n, pbf_int3 dims,
                              uint32_t* tileEnd, const SceneTable& st, void* stream) {
    if (n == 0) return PBF_OK;
    const uint32_t gx = static_cast<uint32_t>(dims.x);
    const uint32_t chunks = (gx + kSweepTileW - 1) / kSweepTileW;
    if (st.geom != nullptr)
        PBF_LAUNCH(k_cell_ranges<true>, grid_for(n), kBlock, stream, sk, cs, ce, n, gx, chunks,
                   tileEnd, st);
    PBF_LAUNCH(k_cell_ranges<false>, grid_for(n), kBlock, stream, sk, cs, ce, n, gx, chunks,
               tileEnd, st);
}

pbf_status launch_ranges_gather(const uint32_t* sk, int* cs, int* ce, size_t n, pbf_int3 dims,
                                uint32_t* tileEnd, const SceneTable& st, const Permute& pm,
                                void* stream) {
    if (n == 0) return PBF_OK;
    if (pm.x == nullptr && pm.permOut == nullptr)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "a p*-only permutation needs permOut");
    if (pm.permOut != nullptr && pm.permOut == pm.permIn)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "permutation composed in place");
    const uint32_t gx = static_cast<uint32_t>(dims.x);
    const uint32_t chunks = (gx + kSweepTileW - 1) / kSweepTileW;
    if (st.geom != nullptr)
        PBF_LAUNCH(k_ranges_gather<true>, grid_for(n), kBlock, stream, sk, cs, ce, n, gx, chunks,
                   tileEnd, st, pm);
    PBF_LAUNCH(k_ranges_gather<false>, grid_for(n), kBlock, stream, sk, cs, ce, n, gx, chunks,
               tileEnd, st, pm);
}

size_t radix_scratch_words(size_t n) {
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    return tiles * kRadix + kRadix + 64;
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
