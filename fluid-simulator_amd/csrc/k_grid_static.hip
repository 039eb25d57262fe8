// pbf_grid_build_static: SpatialGrid::build (src/pbf/spatial_grid.cu:139-226)
// into caller-owned outputs with no allocation, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).
//
// pbf_grid_build runs on a grid handle whose initialize allocated the
// unsorted keys, the sort's ping-pong buffers and the radix counters.  This
// variant keeps those four scratch buffers in one file-scope device array of
// fixed size instead.  The caller passes the inputs (predicted positions,
// bounds, cell size) and the outputs (sorted keys, sorted indices,
// cellStart, cellEnd); the kernels are the ones pbf_grid_build launches.
//
// Each GPU holds its own copy of the array for the life of the process, at
// its own address: the address is taken once per device with
// hipGetSymbolAddress on the first call made with that device current, and
// only that address is passed to kernels (the array's name is a host-side
// stand-in on the host).  Calls on one GPU share the array, so calls that
// overlap on one GPU (two streams, two threads) race; calls in stream order
// do not.  Nothing needs clearing between calls: every pass writes the
// scratch words it later reads.
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "host_internal.hpp"
#include "pbf_device.hpp"
#include "pbf_internal.hpp"
#include "pbf_launch.hpp"

namespace {

constexpr size_t kMaxParticles = PBF_GRID_BUILD_STATIC_MAX_PARTICLES;
// Words of each scratch region: unsorted keys, key ping-pong, value
// ping-pong, then the radix counters and scan partials.
constexpr size_t kKeyWords = kMaxParticles;
constexpr size_t kSortTile = 1024;  // keys per radix tile (radix_scratch_words)
constexpr size_t kCounterWords = (kMaxParticles + kSortTile - 1) / kSortTile * 256 + 256 + 64;
constexpr size_t kScratchWords = 3 * kKeyWords + kCounterWords;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

std::mutex g_mutex;
std::vector<uint32_t*> g_scratch;  // per device: the array's address, or null

// The array's device address on the current device, taken once per device.
pbf_status scratch_address(uint32_t** out) {
    int device = 0;
    pbf_status st = pbf_hip_check(hipGetDevice(&device));
    if (st != PBF_OK) return st;
    if (device < 0) return pbf_fail(PBF_ERR_DEVICE, "no current device");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch.size() <= static_cast<size_t>(device)) g_scratch.resize(device + 1, nullptr);
    if (g_scratch[device] == nullptr) {
        void* p = nullptr;
        st = pbf_hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (st != PBF_OK) return st;
        if (p == nullptr) return pbf_fail(PBF_ERR_DEVICE, "scratch array has no device address");
        g_scratch[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[device];
    return PBF_OK;
}

}  // namespace

extern "C" pbf_status pbf_grid_build_static(const pbf_float4* d_predicted, size_t n,
                                            pbf_float3 minBounds, pbf_float3 maxBounds,
                                            float cellSize, uint32_t* d_sortedKeys,
                                            uint32_t* d_sortedIndices, int* d_cellStart,
                                            int* d_cellEnd, size_t cellCapacity, void* stream) {
    // The checks of SpatialGrid::initialize and build (spatial_grid.cu:11-27,
    // :141-150), all on the host before any device call.
    if (!std::isfinite(cellSize) || cellSize <= 0.0f)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "cellSize must be greater than zero");
    if (!std::isfinite(minBounds.x) || !std::isfinite(minBounds.y) ||
        !std::isfinite(minBounds.z) || !std::isfinite(maxBounds.x) ||
        !std::isfinite(maxBounds.y) || !std::isfinite(maxBounds.z) ||
        maxBounds.x <= minBounds.x || maxBounds.y <= minBounds.y || maxBounds.z <= minBounds.z)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "Invalid spatial grid bounds");
    size_t cells = 0;
    pbf_int3 dims{};
    try {
        const auto d = pbf::GridCore::dimsFor(minBounds, maxBounds, cellSize, &cells);
        dims = pbf_int3{d.x, d.y, d.z};
    } catch (const std::length_error& e) {
        return pbf_fail(PBF_ERR_LENGTH, e.what());
    }
    if (n > kMaxParticles)
        return pbf_fail(PBF_ERR_LENGTH,
                        "pbf_grid_build_static takes at most "
                        "PBF_GRID_BUILD_STATIC_MAX_PARTICLES (1048576) particles");
    if (d_cellStart == nullptr || d_cellEnd == nullptr)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "cell range outputs must not be null");
    if (cellCapacity < cells)
        return pbf_fail(PBF_ERR_OUT_OF_RANGE, "cell range outputs hold fewer cells than the grid");
    if (n != 0 && (d_predicted == nullptr || d_sortedKeys == nullptr || d_sortedIndices == nullptr))
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "predictedPositions and outputs must not be null");
    if (n != 0 && static_cast<const void*>(d_sortedKeys) == static_cast<const void*>(d_sortedIndices))
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "sorted keys and indices must not alias");

    if (n == 0)  // an empty build still clears the ranges (spatial_grid.cu:152-161)
        return pbfint::launch_fill_cells(d_cellStart, d_cellEnd, cells, stream);

    uint32_t* base = nullptr;
    pbf_status st = scratch_address(&base);
    if (st != PBF_OK) return st;
    uint32_t* keys = base;
    uint32_t* tmpK = base + kKeyWords;
    uint32_t* tmpV = base + 2 * kKeyWords;
    uint32_t* counters = base + 3 * kKeyWords;
    if (pbfint::radix_scratch_words(n) > kCounterWords)
        return pbf_fail(PBF_ERR_LENGTH, "radix counters exceed the static scratch");

    st = pbfint::launch_keys(d_predicted, keys, n, minBounds, cellSize, dims,
                             pbfint::SceneTable{}, stream);
    if (st != PBF_OK) return st;
    st = pbfint::radix_sort_pairs(keys, nullptr, d_sortedKeys, d_sortedIndices, tmpK, tmpV, n,
                                  pbfint::radix_key_bits(cells), counters, stream);
    if (st != PBF_OK) return st;
    st = pbfint::launch_fill_cells(d_cellStart, d_cellEnd, cells, stream);
    if (st != PBF_OK) return st;
    return pbfint::launch_cell_ranges(d_sortedKeys, d_cellStart, d_cellEnd, n, stream);
}
