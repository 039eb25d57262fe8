// Library-internal host state shared by SpatialGrid, CollisionSystem and
// PBFSolver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "pbf/collision_system.hpp"
#include "pbf/device_buffer.hpp"
#include "pbf/spatial_grid.hpp"
#include "pbf/types.hpp"
#include "pbf_colliders.hpp"

namespace pbf {

// Throws the reference exception for a failed launch or HIP call.
inline void launch_check(pbf_status status) { check(status); }
void hip_check(hipError_t error);

// Grid configuration and buffers (SpatialGrid::initialize, spatial_grid.cu:10-97).
struct GridCore {
    float3 minBounds{};
    float3 maxBounds{};
    int3 dims{};
    std::size_t numCells = 0;
    std::size_t maxParticles = 0;
    float cellSize = 0.0f;
    int keyBits = 0;

    DeviceBuffer<std::uint32_t> keys;           // unsorted keys
    DeviceBuffer<std::uint32_t> keysTmp;        // sort ping-pong buffer for keys
    DeviceBuffer<std::uint32_t> indices;        // sort ping-pong buffer for values
    DeviceBuffer<std::uint32_t> sortedKeys;
    DeviceBuffer<std::uint32_t> sortedIndices;
    DeviceBuffer<int> cellStart;
    DeviceBuffer<int> cellEnd;
    DeviceBuffer<std::uint32_t> scratch;        // sort histograms, scan partials, re-sort counters
    // The last sort's keys at multiples of the re-sort tile, written with
    // its cell ranges.
    DeviceBuffer<std::uint32_t> splitters;
    // Particle count of the last sort when the caller's particles are in
    // its output order (0: not known to be): a sort of that many keys may
    // then re-sort from it.
    std::size_t orderedCount = 0;

    // Validates like the reference and allocates; throws on failure.
    static GridCore create(std::size_t maxParticles, float3 minBounds, float3 maxBounds,
                           float cellSize);
    // Cells per axis over [minBounds, maxBounds] and their count, validated
    // like create (which uses it).
    static int3 dimsFor(float3 minBounds, float3 maxBounds, float cellSize, std::size_t* cells);
    // The keyed part of a grid for a batch of scenes: numCells global cells
    // (the scenes' grids laid end to end), no bounds of its own.
    static GridCore createKeyed(std::size_t maxParticles, std::size_t numCells, float cellSize);
    // scratch (zeroed) and splitters for maxParticles keys.
    void allocateSortScratch();

    // Stable sort of `keys` (slot values implicit) into sortedKeys and
    // sortedIndices: re-sorted from the last sort when `resort` is set and
    // orderedCount == particleCount, else radix-sorted; the results are the
    // same bit for bit.  Leaves orderedCount at 0 (the caller sets it once
    // its particles follow the new order).  Returns whether it re-sorted.
    bool sortKeys(std::size_t particleCount, bool resort, hipStream_t stream);
    // sortKeys, then the cell ranges (and the splitters).
    void sortAndRanges(std::size_t particleCount, bool resort, hipStream_t stream);
};

struct SpatialGrid::State {
    GridCore core;
};

struct CollisionSystem::State {
    DeviceBuffer<Container> container;
    DeviceBuffer<std::uint32_t> status;  // fixed-point status word
    Container hostContainer{};
    bool hasContainer = false;
    DeviceBuffer<SphereCollider> spheres;
    DeviceBuffer<BoxCollider> boxes;
    DeviceBuffer<PlaneCollider> planes;
    std::size_t sphereCount = 0;
    std::size_t boxCount = 0;
    std::size_t planeCount = 0;

    pbfdev::Colliders view() const {
        return pbfdev::Colliders{container.data(),
                                 spheres.data(),
                                 boxes.data(),
                                 planes.data(),
                                 static_cast<int>(sphereCount),
                                 static_cast<int>(boxCount),
                                 static_cast<int>(planeCount)};
    }
};

}  // namespace pbf
