// Uniform hash grid over the simulation box.
//
// Replaces SpatialGrid (include/pbf/spatial_grid.cuh:11-62,
// src/pbf/spatial_grid.cu:10-268): same validation, dimensions, key layout
// (x + y*gx + z*gx*gy), stable order and -1 marked empty cells.  The sort is
// the library's LDS-staged radix sort instead of CUB.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

#include "pbf/types.hpp"

namespace pbf {

class SpatialGrid {
public:
    SpatialGrid();
    ~SpatialGrid();
    SpatialGrid(SpatialGrid&&) noexcept;
    SpatialGrid& operator=(SpatialGrid&&) noexcept;
    SpatialGrid(const SpatialGrid&) = delete;
    SpatialGrid& operator=(const SpatialGrid&) = delete;

    void initialize(std::size_t maxParticles, float3 minBounds, float3 maxBounds, float cellSize);

    // Keys, stable sort and cell ranges of device positions; synchronises.
    void build(const float4* predictedPositions, std::size_t particleCount);
    // build's results bit for bit, computed by re-sorting from the last
    // build: fast when the positions are that build's particles laid out in
    // its sorted order (slot i holding particle sortedIndices()[i]) after a
    // small move, so that few cell keys changed.  Without a previous build
    // of particleCount particles it is build.  No reference counterpart.
    void rebuild(const float4* predictedPositions, std::size_t particleCount);

    const std::uint32_t* sortedKeys() const;
    const std::uint32_t* sortedIndices() const;
    const int* cellStart() const;
    const int* cellEnd() const;

    int3 gridSize() const;
    std::size_t numCells() const;
    std::size_t maxParticles() const;
    float cellSize() const;
    float3 minBounds() const;
    float3 maxBounds() const;

    struct State;  // library-internal

private:
    void build_impl(const float4* predictedPositions, std::size_t particleCount, bool resort);
    std::unique_ptr<State> state_;
};

}  // namespace pbf
