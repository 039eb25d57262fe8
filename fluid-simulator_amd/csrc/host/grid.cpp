// SpatialGrid and the grid core the solver shares.
//
// Reference: src/pbf/spatial_grid.cu:10-268.
#include <cmath>
#include <limits>
#include <stdexcept>

#include "host_internal.hpp"
#include "pbf_internal.hpp"

namespace pbf {

void hip_check(hipError_t error) {
    if (error != hipSuccess) throw std::runtime_error(hipGetErrorString(error));
}

GridCore GridCore::create(std::size_t maxParticles, float3 minBounds, float3 maxBounds,
                          float cellSize) {
    // Validation order and exception types of spatial_grid.cu:11-27.
    if (maxParticles == 0) throw std::invalid_argument("maxParticles must be greater than zero");
    if (!std::isfinite(cellSize) || cellSize <= 0.0f)
        throw std::invalid_argument("cellSize must be greater than zero");
    if (!std::isfinite(minBounds.x) || !std::isfinite(minBounds.y) ||
        !std::isfinite(minBounds.z) || !std::isfinite(maxBounds.x) ||
        !std::isfinite(maxBounds.y) || !std::isfinite(maxBounds.z) ||
        maxBounds.x <= minBounds.x || maxBounds.y <= minBounds.y || maxBounds.z <= minBounds.z) {
        throw std::invalid_argument("Invalid spatial grid bounds");
    }
    if (maxParticles > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::length_error("Particle count exceeds grid range index capacity");

    GridCore next;
    next.maxParticles = maxParticles;
    next.minBounds = minBounds;
    next.maxBounds = maxBounds;
    next.cellSize = cellSize;

    next.dims = dimsFor(minBounds, maxBounds, cellSize, &next.numCells);
    next.keyBits = pbfint::radix_key_bits(next.numCells);

    next.keys.allocate(maxParticles);
    next.keysTmp.allocate(maxParticles);
    next.indices.allocate(maxParticles);
    next.sortedKeys.allocate(maxParticles);
    next.sortedIndices.allocate(maxParticles);
    next.cellStart.allocate(next.numCells);
    next.cellEnd.allocate(next.numCells);
    next.allocateSortScratch();
    return next;
}

void GridCore::allocateSortScratch() {
    scratch.allocate(pbfint::grid_scratch_words(maxParticles));
    scratch.fillBytes(0);  // the re-sort counters start at zero
    // The memset runs on the null stream, which the solver's non-blocking
    // stream does not wait for.
    hip_check(hipDeviceSynchronize());
    splitters.allocate(pbfint::resort_splitter_count(maxParticles));
    orderedCount = 0;
}

int3 GridCore::dimsFor(float3 minBounds, float3 maxBounds, float cellSize, std::size_t* cells) {
    // ceil((max - min) / cellSize) in double (spatial_grid.cu:37-53).
    const auto axis = [cellSize](float lo, float hi) {
        const double count = std::ceil((static_cast<double>(hi) - static_cast<double>(lo)) /
                                       static_cast<double>(cellSize));
        if (!std::isfinite(count) || count < 1.0 ||
            count > static_cast<double>(std::numeric_limits<int>::max()))
            throw std::length_error("Spatial grid axis exceeds integer capacity");
        return static_cast<int>(count);
    };
    int3 d;
    d.x = axis(minBounds.x, maxBounds.x);
    d.y = axis(minBounds.y, maxBounds.y);
    d.z = axis(minBounds.z, maxBounds.z);
    const std::size_t maxCells = static_cast<std::size_t>(std::numeric_limits<std::uint32_t>::max());
    const std::size_t cx = static_cast<std::size_t>(d.x);
    const std::size_t cy = static_cast<std::size_t>(d.y);
    const std::size_t cz = static_cast<std::size_t>(d.z);
    if (cx > maxCells / cy || cx * cy > maxCells / cz)
        throw std::length_error("Spatial grid exceeds 32-bit key capacity");
    *cells = cx * cy * cz;
    return d;
}

GridCore GridCore::createKeyed(std::size_t maxParticles, std::size_t numCells, float cellSize) {
    if (numCells == 0 || numCells > static_cast<std::size_t>(std::numeric_limits<std::uint32_t>::max()))
        throw std::length_error("Batched grids exceed 32-bit key capacity");
    GridCore next;
    next.maxParticles = maxParticles;
    next.cellSize = cellSize;
    next.numCells = numCells;
    next.keyBits = pbfint::radix_key_bits(numCells);
    next.keys.allocate(maxParticles);
    next.keysTmp.allocate(maxParticles);
    next.indices.allocate(maxParticles);
    next.sortedKeys.allocate(maxParticles);
    next.sortedIndices.allocate(maxParticles);
    next.cellStart.allocate(numCells);
    next.cellEnd.allocate(numCells);
    next.allocateSortScratch();
    return next;
}

bool GridCore::sortKeys(std::size_t n, bool resort, hipStream_t stream) {
    if (resort && n != 0 && orderedCount == n) {
        orderedCount = 0;  // until the caller's particles follow this sort
        launch_check(pbfint::resort_pairs(keys.data(), sortedKeys.data(), splitters.data(),
                                          sortedKeys.data(), sortedIndices.data(), keysTmp.data(),
                                          indices.data(), n, maxParticles, scratch.data(), stream));
        return true;
    }
    orderedCount = 0;
    launch_check(pbfint::radix_sort_pairs(keys.data(), nullptr, sortedKeys.data(),
                                          sortedIndices.data(), keysTmp.data(), indices.data(), n,
                                          keyBits, scratch.data(), stream));
    return false;
}

void GridCore::sortAndRanges(std::size_t n, bool resort, hipStream_t stream) {
    sortKeys(n, resort, stream);
    launch_check(pbfint::launch_fill_cells(cellStart.data(), cellEnd.data(), numCells, stream));
    launch_check(pbfint::launch_cell_ranges(sortedKeys.data(), cellStart.data(), cellEnd.data(), n,
                                            stream, splitters.data()));
}

SpatialGrid::SpatialGrid() : state_(new State()) {}
SpatialGrid::~SpatialGrid() = default;
SpatialGrid::SpatialGrid(SpatialGrid&&) noexcept = default;
SpatialGrid& SpatialGrid::operator=(SpatialGrid&&) noexcept = default;

void SpatialGrid::initialize(std::size_t maxParticles, float3 minBounds, float3 maxBounds,
                             float cellSize) {
    GridCore next = GridCore::create(maxParticles, minBounds, maxBounds, cellSize);
    state_->core = std::move(next);
}

void SpatialGrid::build(const float4* predicted, std::size_t n) { build_impl(predicted, n, false); }

void SpatialGrid::rebuild(const float4* predicted, std::size_t n) { build_impl(predicted, n, true); }

void SpatialGrid::build_impl(const float4* predicted, std::size_t n, bool resort) {
    GridCore& g = state_->core;
    if (g.maxParticles == 0) throw std::logic_error("SpatialGrid must be initialized before building");
    if (predicted == nullptr && n != 0)
        throw std::invalid_argument("predictedPositions must not be null");
    if (n > g.maxParticles)
        throw std::out_of_range("Particle count exceeds maximum number of particles");
    hipStream_t stream = nullptr;
    if (n == 0) {
        // An empty build still clears the ranges of a previous build.
        launch_check(pbfint::launch_fill_cells(g.cellStart.data(), g.cellEnd.data(), g.numCells,
                                               stream));
        g.orderedCount = 0;
        hip_check(hipDeviceSynchronize());
        return;
    }
    launch_check(pbfint::launch_keys(predicted, g.keys.data(), n, g.minBounds, g.cellSize, g.dims,
                                     pbfint::SceneTable{}, stream));
    g.sortAndRanges(n, resort, stream);
    // Surface asynchronous failures from this call (spatial_grid.cu:220-225).
    hip_check(hipDeviceSynchronize());
    // Positions laid out in this build's order may be re-sorted from it.
    g.orderedCount = n;
}

const std::uint32_t* SpatialGrid::sortedKeys() const { return state_->core.sortedKeys.data(); }
const std::uint32_t* SpatialGrid::sortedIndices() const { return state_->core.sortedIndices.data(); }
const int* SpatialGrid::cellStart() const { return state_->core.cellStart.data(); }
const int* SpatialGrid::cellEnd() const { return state_->core.cellEnd.data(); }
int3 SpatialGrid::gridSize() const { return state_->core.dims; }
std::size_t SpatialGrid::numCells() const { return state_->core.numCells; }
std::size_t SpatialGrid::maxParticles() const { return state_->core.maxParticles; }
float SpatialGrid::cellSize() const { return state_->core.cellSize; }
float3 SpatialGrid::minBounds() const { return state_->core.minBounds; }
float3 SpatialGrid::maxBounds() const { return state_->core.maxBounds; }

}  // namespace pbf
