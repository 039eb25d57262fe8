// C ABI over the C++ classes: opaque handles, exceptions -> status codes.
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "host_internal.hpp"
#include "pbf/pbf_adjoint.hpp"
#include "pbf/pbf_batch.hpp"
#include "pbf/pbf_solver.hpp"
#include "pbf_amd.h"
#include "pbf_launch.hpp"

struct pbf_grid {
    pbf::SpatialGrid grid;
};
struct pbf_collision {
    pbf::CollisionSystem system;
};
struct pbf_solver {
    pbf::PBFSolver solver;
};
struct pbf_batch {
    pbf::PBFBatchSolver batch;
};
struct pbf_adjoint {
    explicit pbf_adjoint(pbf::PBFSolver& s) : adjoint(s) {}
    pbf::PBFAdjoint adjoint;
};

namespace {

// Runs f and maps the reference's exception types onto status codes.
template <typename F>
pbf_status guarded(F&& f) {
    try {
        f();
        return PBF_OK;
    } catch (const pbf::collision_nonconvergence& e) {
        return pbf_fail(PBF_ERR_NONCONVERGENCE, e.what());
    } catch (const std::overflow_error& e) {
        return pbf_fail(PBF_ERR_OVERFLOW, e.what());
    } catch (const std::invalid_argument& e) {
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::out_of_range& e) {
        return pbf_fail(PBF_ERR_OUT_OF_RANGE, e.what());
    } catch (const std::length_error& e) {
        return pbf_fail(PBF_ERR_LENGTH, e.what());
    } catch (const std::logic_error& e) {
        return pbf_fail(PBF_ERR_LOGIC, e.what());
    } catch (const std::bad_alloc& e) {
        return pbf_fail(PBF_ERR_DEVICE, e.what());
    } catch (const std::exception& e) {
        return pbf_fail(PBF_ERR_DEVICE, e.what());
    } catch (...) {
        return pbf_fail(PBF_ERR_DEVICE, "unknown error");
    }
}

#define PBF_REQUIRE(ptr)                                                         \
    do {                                                                         \
        if ((ptr) == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null handle"); \
    } while (0)

template <typename T>
std::vector<T> to_vector(const T* data, size_t count) {
    if (count != 0 && data == nullptr) throw std::invalid_argument("null collider array");
    return std::vector<T>(data, data + count);
}

}  // namespace

namespace {

void write_kernel_times(const std::vector<pbf::PBFSolver::KernelTime>& times, char* names,
                        size_t namesBytes, double* ms, uint64_t* launches, size_t capacity,
                        size_t* kernelCount) {
    if (kernelCount) *kernelCount = times.size();
    std::string joined;
    for (size_t i = 0; i < times.size(); ++i) {
        if (i < capacity) {
            if (ms) ms[i] = times[i].milliseconds;
            if (launches) launches[i] = times[i].launches;
        }
        joined += times[i].name;
        joined += '\n';
    }
    if (names != nullptr && namesBytes > 0) {
        const size_t len = joined.size() < namesBytes - 1 ? joined.size() : namesBytes - 1;
        std::memcpy(names, joined.data(), len);
        names[len] = '\0';
    }
}

}  // namespace

namespace {
pbf::PBFSolver::Tuning from_c(const pbf_tuning& t) {
    return pbf::PBFSolver::Tuning{t.neighborSkin, t.passBlock, t.particlesPerThread};
}
pbf_tuning to_c(const pbf::PBFSolver::Tuning& t) {
    return pbf_tuning{t.neighborSkin, t.passBlock, t.particlesPerThread};
}

template <class Tuner>
pbf_status run_autotune(Tuner&& tune, const pbf_tuning* candidates, size_t count,
                        double* msPerFrame, pbf_tuning* chosen) {
    if (candidates == nullptr && count != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "candidates must not be null");
    return guarded([&] {
        std::vector<pbf::PBFSolver::Tuning> c;
        for (size_t k = 0; k < count; ++k) c.push_back(from_c(candidates[k]));
        const std::vector<pbf::PBFSolver::TuningTime> times = tune(c);
        if (msPerFrame != nullptr)
            for (size_t k = 0; k < times.size(); ++k) msPerFrame[k] = times[k].msPerFrame;
        if (chosen != nullptr) {
            const pbf::PBFSolver::TuningTime* best = nullptr;
            for (const auto& t : times)
                if (best == nullptr || t.msPerFrame < best->msPerFrame) best = &t;
            if (best != nullptr) *chosen = to_c(best->tuning);
        }
    });
}
}  // namespace

extern "C" {

// ---------------------------------------------------------------- grid
pbf_status pbf_grid_create(pbf_grid** grid) {
    PBF_REQUIRE(grid);
    return guarded([&] { *grid = new pbf_grid(); });
}

pbf_status pbf_grid_destroy(pbf_grid* grid) {
    delete grid;
    return PBF_OK;
}

pbf_status pbf_grid_initialize(pbf_grid* grid, size_t maxParticles, pbf_float3 minBounds,
                               pbf_float3 maxBounds, float cellSize) {
    PBF_REQUIRE(grid);
    return guarded([&] { grid->grid.initialize(maxParticles, minBounds, maxBounds, cellSize); });
}

pbf_status pbf_grid_build(pbf_grid* grid, const pbf_float4* d_predicted, size_t n) {
    PBF_REQUIRE(grid);
    return guarded([&] { grid->grid.build(d_predicted, n); });
}

pbf_status pbf_grid_rebuild(pbf_grid* grid, const pbf_float4* d_predicted, size_t n) {
    PBF_REQUIRE(grid);
    return guarded([&] { grid->grid.rebuild(d_predicted, n); });
}

const uint32_t* pbf_grid_sorted_keys(const pbf_grid* g) { return g ? g->grid.sortedKeys() : nullptr; }
const uint32_t* pbf_grid_sorted_indices(const pbf_grid* g) {
    return g ? g->grid.sortedIndices() : nullptr;
}
const int* pbf_grid_cell_start(const pbf_grid* g) { return g ? g->grid.cellStart() : nullptr; }
const int* pbf_grid_cell_end(const pbf_grid* g) { return g ? g->grid.cellEnd() : nullptr; }

pbf_status pbf_grid_info(const pbf_grid* g, pbf_int3* gridSize, size_t* numCells,
                         size_t* maxParticles, float* cellSize, pbf_float3* minBounds,
                         pbf_float3* maxBounds) {
    PBF_REQUIRE(g);
    if (gridSize) *gridSize = g->grid.gridSize();
    if (numCells) *numCells = g->grid.numCells();
    if (maxParticles) *maxParticles = g->grid.maxParticles();
    if (cellSize) *cellSize = g->grid.cellSize();
    if (minBounds) *minBounds = g->grid.minBounds();
    if (maxBounds) *maxBounds = g->grid.maxBounds();
    return PBF_OK;
}

// ---------------------------------------------------------------- collisions
pbf_status pbf_collision_create(pbf_collision** cs) {
    PBF_REQUIRE(cs);
    return guarded([&] { *cs = new pbf_collision(); });
}

pbf_status pbf_collision_destroy(pbf_collision* cs) {
    delete cs;
    return PBF_OK;
}

pbf_status pbf_collision_set_container(pbf_collision* cs, pbf_container c) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.setContainer(c); });
}

pbf_status pbf_collision_set_spheres(pbf_collision* cs, const pbf_sphere* s, size_t n) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.setSpheres(to_vector(s, n)); });
}
pbf_status pbf_collision_clear_spheres(pbf_collision* cs) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.clearSpheres(); });
}
pbf_status pbf_collision_set_boxes(pbf_collision* cs, const pbf_box* b, size_t n) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.setBoxes(to_vector(b, n)); });
}
pbf_status pbf_collision_clear_boxes(pbf_collision* cs) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.clearBoxes(); });
}
pbf_status pbf_collision_set_planes(pbf_collision* cs, const pbf_plane* p, size_t n) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.setPlanes(to_vector(p, n)); });
}
pbf_status pbf_collision_clear_planes(pbf_collision* cs) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.clearPlanes(); });
}

pbf_status pbf_collision_solve(pbf_collision* cs, pbf_float4* d_predicted, size_t n, float r) {
    PBF_REQUIRE(cs);
    return guarded([&] { cs->system.solve(d_predicted, n, r); });
}

pbf_status pbf_collision_resolve_velocities(pbf_collision* cs, const pbf_float4* d_position,
                                            const pbf_float4* d_incoming, pbf_float4* d_velocity,
                                            size_t n, float r, float e, float f) {
    PBF_REQUIRE(cs);
    return guarded([&] {
        cs->system.resolveVelocities(d_position, d_incoming, d_velocity, n, r, e, f);
        pbf::hip_check(hipDeviceSynchronize());
    });
}

// ---------------------------------------------------------------- solver
pbf_status pbf_solver_create(pbf_solver** solver) {
    PBF_REQUIRE(solver);
    return guarded([&] { *solver = new pbf_solver(); });
}

pbf_status pbf_solver_destroy(pbf_solver* solver) {
    delete solver;
    return PBF_OK;
}

pbf_status pbf_solver_initialize(pbf_solver* s, size_t maxParticles, pbf_float3 minBounds,
                                 pbf_float3 maxBounds, const pbf_sim_params* params) {
    PBF_REQUIRE(s);
    if (params == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null params");
    return guarded([&] { s->solver.initialize(maxParticles, minBounds, maxBounds, *params); });
}

pbf_status pbf_solver_set_particles(pbf_solver* s, const pbf_float4* pos, const pbf_float4* vel,
                                    size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setParticles(pos, vel, n); });
}

pbf_status pbf_solver_set_particles_device(pbf_solver* s, const pbf_float4* pos,
                                           const pbf_float4* vel, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setParticlesDevice(pos, vel, n); });
}

pbf_status pbf_solver_set_spheres(pbf_solver* s, const pbf_sphere* sp, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setSpheres(to_vector(sp, n)); });
}
pbf_status pbf_solver_set_boxes(pbf_solver* s, const pbf_box* b, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setBoxes(to_vector(b, n)); });
}
pbf_status pbf_solver_set_planes(pbf_solver* s, const pbf_plane* p, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setPlanes(to_vector(p, n)); });
}

pbf_status pbf_solver_step(pbf_solver* s) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.step(); });
}

pbf_status pbf_solver_run_one(pbf_solver* s) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.run(); });
}

pbf_status pbf_solver_run(pbf_solver* s, size_t frames) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.run(frames); });
}

pbf_status pbf_solver_copy_positions_to_host(pbf_solver* s, pbf_float4* out, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.copyPositionsToHost(out, n); });
}
pbf_status pbf_solver_copy_velocities_to_host(pbf_solver* s, pbf_float4* out, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.copyVelocitiesToHost(out, n); });
}
pbf_status pbf_solver_copy_positions_to_device(pbf_solver* s, pbf_float4* out, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.copyPositionsToDevice(out, n); });
}
pbf_status pbf_solver_copy_velocities_to_device(pbf_solver* s, pbf_float4* out, size_t n) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.copyVelocitiesToDevice(out, n); });
}

pbf_status pbf_solver_counts(const pbf_solver* s, size_t* count, size_t* maxParticles) {
    PBF_REQUIRE(s);
    if (count) *count = s->solver.particleCount();
    if (maxParticles) *maxParticles = s->solver.maxParticles();
    return PBF_OK;
}

void* pbf_solver_stream(pbf_solver* s) { return s ? s->solver.stream() : nullptr; }

pbf_status pbf_solver_set_graphs(pbf_solver* s, int enabled) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setGraphs(enabled != 0); });
}

pbf_status pbf_solver_set_incremental_sort(pbf_solver* s, int enabled) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setIncrementalSort(enabled != 0); });
}

pbf_status pbf_solver_set_fused_xsph(pbf_solver* s, int enabled) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setFusedXsph(enabled != 0); });
}

pbf_status pbf_solver_set_neighbor_skin(pbf_solver* s, float skin) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setNeighborSkin(skin); });
}

pbf_status pbf_solver_neighbor_skin(const pbf_solver* s, float* skin, size_t* staleFrames) {
    PBF_REQUIRE(s);
    if (skin) *skin = s->solver.neighborSkin();
    if (staleFrames) *staleFrames = s->solver.staleFrames();
    return PBF_OK;
}


pbf_status pbf_solver_set_tuning(pbf_solver* s, const pbf_tuning* t) {
    PBF_REQUIRE(s);
    PBF_REQUIRE(t);
    return guarded([&] { s->solver.setTuning(from_c(*t)); });
}

pbf_status pbf_solver_get_tuning(const pbf_solver* s, pbf_tuning* t) {
    PBF_REQUIRE(s);
    PBF_REQUIRE(t);
    *t = to_c(s->solver.tuning());
    return PBF_OK;
}

pbf_status pbf_solver_autotune(pbf_solver* s, size_t frames, const pbf_tuning* candidates,
                               size_t count, double* msPerFrame, pbf_tuning* chosen) {
    PBF_REQUIRE(s);
    return run_autotune(
        [&](const std::vector<pbf::PBFSolver::Tuning>& c) { return s->solver.autotune(frames, c); },
        candidates, count, msPerFrame, chosen);
}

pbf_status pbf_solver_set_profiling(pbf_solver* s, int enabled) {
    PBF_REQUIRE(s);
    return guarded([&] { s->solver.setProfiling(enabled != 0); });
}

pbf_status pbf_solver_kernel_times(pbf_solver* s, char* names, size_t namesBytes, double* ms,
                                   uint64_t* launches, size_t capacity, size_t* kernelCount) {
    PBF_REQUIRE(s);
    return guarded([&] {
        write_kernel_times(s->solver.kernelTimes(), names, namesBytes, ms, launches, capacity,
                           kernelCount);
    });
}

pbf_status pbf_solver_neighbor_lists(pbf_solver* s, pbf_float4* positions, int* counts,
                                     uint32_t* neighbors, size_t n, int maxNeighbors) {
    PBF_REQUIRE(s);
    return guarded([&] {
        s->solver.copyNeighborLists(positions, counts, neighbors, n, maxNeighbors);
    });
}

pbf_status pbf_batch_create(pbf_batch** batch) {
    if (batch == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *batch = new pbf_batch(); });
}

pbf_status pbf_batch_destroy(pbf_batch* batch) {
    delete batch;
    return PBF_OK;
}

pbf_status pbf_batch_initialize(pbf_batch* b, const pbf_batch_scene* scenes, size_t sceneCount,
                                const pbf_sim_params* params) {
    PBF_REQUIRE(b);
    if (params == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null parameters");
    if (scenes == nullptr && sceneCount != 0)
        return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null scene table");
    return guarded([&] {
        std::vector<pbf::BatchScene> v(sceneCount);
        for (size_t k = 0; k < sceneCount; ++k)
            v[k] = pbf::BatchScene{scenes[k].particleCount, scenes[k].minBounds, scenes[k].maxBounds};
        b->batch.initialize(v, *params);
    });
}

pbf_status pbf_batch_set_particles(pbf_batch* b, const pbf_float4* pos, const pbf_float4* vel,
                                   size_t n) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setParticles(pos, vel, n); });
}

pbf_status pbf_batch_step(pbf_batch* b) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.step(); });
}

pbf_status pbf_batch_run(pbf_batch* b, size_t frames) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.run(frames); });
}

pbf_status pbf_batch_copy_positions_to_host(pbf_batch* b, pbf_float4* out, size_t n) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.copyPositionsToHost(out, n); });
}

pbf_status pbf_batch_copy_velocities_to_host(pbf_batch* b, pbf_float4* out, size_t n) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.copyVelocitiesToHost(out, n); });
}

pbf_status pbf_batch_counts(const pbf_batch* b, size_t* sceneCount, size_t* particleCount) {
    PBF_REQUIRE(b);
    if (sceneCount) *sceneCount = b->batch.sceneCount();
    if (particleCount) *particleCount = b->batch.particleCount();
    return PBF_OK;
}

pbf_status pbf_batch_set_graphs(pbf_batch* b, int enabled) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setGraphs(enabled != 0); });
}

pbf_status pbf_batch_set_incremental_sort(pbf_batch* b, int enabled) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setIncrementalSort(enabled != 0); });
}

pbf_status pbf_batch_set_fused_xsph(pbf_batch* b, int enabled) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setFusedXsph(enabled != 0); });
}

pbf_status pbf_batch_set_neighbor_skin(pbf_batch* b, float skin) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setNeighborSkin(skin); });
}

pbf_status pbf_batch_stale_frames(const pbf_batch* b, size_t* staleFrames) {
    PBF_REQUIRE(b);
    if (staleFrames) *staleFrames = b->batch.staleFrames();
    return PBF_OK;
}

pbf_status pbf_batch_set_tuning(pbf_batch* b, const pbf_tuning* t) {
    PBF_REQUIRE(b);
    PBF_REQUIRE(t);
    return guarded([&] { b->batch.setTuning(from_c(*t)); });
}

pbf_status pbf_batch_autotune(pbf_batch* b, size_t frames, const pbf_tuning* candidates,
                              size_t count, double* msPerFrame, pbf_tuning* chosen) {
    PBF_REQUIRE(b);
    return run_autotune(
        [&](const std::vector<pbf::PBFSolver::Tuning>& c) { return b->batch.autotune(frames, c); },
        candidates, count, msPerFrame, chosen);
}

pbf_status pbf_batch_set_profiling(pbf_batch* b, int enabled) {
    PBF_REQUIRE(b);
    return guarded([&] { b->batch.setProfiling(enabled != 0); });
}

pbf_status pbf_batch_kernel_times(pbf_batch* b, char* names, size_t namesBytes, double* ms,
                                  uint64_t* launches, size_t capacity, size_t* kernelCount) {
    PBF_REQUIRE(b);
    return guarded([&] {
        write_kernel_times(b->batch.kernelTimes(), names, namesBytes, ms, launches, capacity,
                           kernelCount);
    });
}

pbf_status pbf_adjoint_create(pbf_solver* s, pbf_adjoint** adjoint) {
    PBF_REQUIRE(s);
    if (adjoint == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *adjoint = new pbf_adjoint(s->solver); });
}

pbf_status pbf_adjoint_destroy(pbf_adjoint* adjoint) {
    delete adjoint;
    return PBF_OK;
}

pbf_status pbf_adjoint_forward(pbf_adjoint* a, size_t frames, size_t every) {
    PBF_REQUIRE(a);
    return guarded([&] { a->adjoint.forward(frames, every); });
}

pbf_status pbf_adjoint_backward(pbf_adjoint* a, const pbf_float4* xb, const pbf_float4* vb,
                                size_t n, pbf_float4* x0b, pbf_float4* v0b, double* cb,
                                double* eb) {
    PBF_REQUIRE(a);
    return guarded([&] { a->adjoint.backward(xb, vb, n, x0b, v0b, cb, eb); });
}

}  // extern "C"
