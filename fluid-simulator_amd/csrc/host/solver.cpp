// PBFSolver: owns the working set in HBM and enqueues the fused step.
//
// Reference: src/pbf/pbf_solver.cu:1-379.
//
// The working set stays in cell order between steps, as in the reference;
// stable ids map it back to API order on copy-out.  Per substep the device
// runs (reference stage -> kernel here):
//   integrate + v_in snapshot + first cell keys      -> k_predict_key
//   per iteration: grid build, reorder               -> re-sort (or radix sort),
//                                                         k_ranges_gather
//                  neighbour lists                    -> k_build_lists (LDS cell sweep)
//                  lambda (into p*.w)                 -> k_lambda_list
//                  delta-p, apply, collision solve    -> k_delta_apply_collide
//                  (restores p*.w from x.w; writes the next build's keys)
//   post-solve rebuild (if XSPH or vorticity)         -> as above
//   velocity update + XSPH                            -> in the post-solve
//                                                         k_build_lists, or
//                                                         k_update_xsph
//   vorticity (+ (x, |omega|) records)                -> k_vorticity_list
//   confinement + velocity contacts                   -> k_confine_resolve
// The lists hold the reference's neighbours in the reference's order, stored
// in groups of four entries (one 16-byte access per four neighbours) rather
// than slot-major; copyNeighborLists hands them out slot-major.  The list
// kernels run their fast form (hardware rsq in the spiky gradient,
// workgroups contiguous per XCD).
// Neighbour overflow and collision non-convergence set bits of a device
// status word; every later kernel that changes the state returns at once,
// and the host reads the word once when the step (or run) finishes.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "host_internal.hpp"
#include "pbf/pbf_adjoint.hpp"
#include "pbf/pbf_batch.hpp"
#include "pbf/pbf_solver.hpp"
#include "pbf_internal.hpp"

namespace pbf {
namespace {

constexpr int kBlockSize = 256;     // pbf_solver.cu:11
constexpr int kMaxNeighbors = 256;  // pbf_solver.cu:12

// HIP-event timing of named launches on one stream.
class Profiler {
public:
    ~Profiler() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (const Pending& p : pending_) {
            (void)hipEventDestroy(p.begin);
            (void)hipEventDestroy(p.end);
        }
    }

    bool enabled = false;

    void begin(const char* name, hipStream_t stream) {
        if (!enabled) return;
        Pending p;
        p.id = id_of(name);
        p.begin = take();
        p.end = take();
        hip_check(hipEventRecord(p.begin, stream));
        pending_.push_back(p);
    }

    void end(hipStream_t stream) {
        if (!enabled || pending_.empty()) return;
        hip_check(hipEventRecord(pending_.back().end, stream));
    }

    // Call after the stream has been synchronised.
    void collect() {
        for (const Pending& p : pending_) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.begin, p.end) == hipSuccess) {
                ms_[p.id] += ms;
                launches_[p.id] += 1;
            }
            pool_.push_back(p.begin);
            pool_.push_back(p.end);
        }
        pending_.clear();
    }

    std::vector<PBFSolver::KernelTime> times() const {
        std::vector<PBFSolver::KernelTime> out;
        for (std::size_t i = 0; i < names_.size(); ++i)
            out.push_back({names_[i], ms_[i], launches_[i]});
        return out;
    }

    void reset() {
        for (auto& m : ms_) m = 0.0;
        for (auto& l : launches_) l = 0;
    }

private:
    struct Pending {
        int id;
        hipEvent_t begin;
        hipEvent_t end;
    };

    int id_of(const char* name) {
        for (std::size_t i = 0; i < names_.size(); ++i)
            if (names_[i] == name) return static_cast<int>(i);
        names_.emplace_back(name);
        ms_.push_back(0.0);
        launches_.push_back(0);
        return static_cast<int>(names_.size() - 1);
    }

    hipEvent_t take() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e));
        return e;
    }

    std::vector<std::string> names_;
    std::vector<double> ms_;
    std::vector<std::uint64_t> launches_;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
};

class Scope {
public:
    Scope(Profiler& p, const char* name, hipStream_t s) : p_(p), s_(s) { p_.begin(name, s_); }
    ~Scope() { p_.end(s_); }

private:
    Profiler& p_;
    hipStream_t s_;
};

bool finite4(const float4& v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}

// Pinned host staging for uploads and copy-outs through pageable memory:
// two chunk buffers used in turn, so the host-side copy (and validation) of
// one chunk overlaps the DMA of the other.  An event per buffer marks the
// end of the last DMA that used it.
class PinnedStage {
public:
    static constexpr std::size_t kChunk = std::size_t{1} << 19;  // float4 per chunk (8 MiB)

    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage() {
        for (int k = 0; k < 2; ++k) {
            if (buf_[k] != nullptr) (void)hipHostFree(buf_[k]);
            if (done_[k] != nullptr) (void)hipEventDestroy(done_[k]);
        }
    }

    float4* buffer(int k) {
        ensure();
        return buf_[k];
    }
    hipEvent_t event(int k) {
        ensure();
        return done_[k];
    }

private:
    void ensure() {
        if (buf_[0] != nullptr) return;
        for (int k = 0; k < 2; ++k) {
            void* p = nullptr;
            hip_check(hipHostMalloc(&p, kChunk * sizeof(float4), hipHostMallocDefault));
            buf_[k] = static_cast<float4*>(p);
            hip_check(hipEventCreateWithFlags(&done_[k], hipEventDisableTiming));
        }
    }

    float4* buf_[2] = {nullptr, nullptr};
    hipEvent_t done_[2] = {nullptr, nullptr};
};

// Where a caller's host-array argument lives: pageable host memory (staged
// through the pinned chunks), memory both the host and the device reach
// (pinned or managed: one DMA), or device memory, which the host-array
// calls reject (the host cannot read or write it).
enum class HostArray { Pageable, Pinned, Device };

HostArray host_array_kind(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory reports an error
        return HostArray::Pageable;
    }
    if (a.type == hipMemoryTypeHost || a.type == hipMemoryTypeManaged) return HostArray::Pinned;
    if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeArray) return HostArray::Device;
    return HostArray::Pageable;
}

}  // namespace

// What the adjoint needs of one substep of the forward, in the working-set
// order of each grid build (one per constraint iteration, then the
// post-solve rebuild at index `iterations`): the particle id of every slot,
// the grouped lists and counts, the positions entering each iteration
// (lambda in .w) and after its Delta-p, before the collision solve (written
// by the Delta-p pass itself), and the post-solve phase's records.  Sized
// for the solver's iteration count.
struct SubstepTape {
    // Every array below lives in one allocation (one device allocation per
    // substep of a segment rather than one per array).
    DeviceBuffer<float4> arena;
    std::size_t n = 0;
    std::size_t listWords = 0;
    std::vector<std::uint32_t*> ids, nbr;
    std::vector<int*> cnt;
    std::vector<float4*> pk, pre;
    float4 *xs = nullptr, *pf = nullptr, *v3 = nullptr, *omega = nullptr, *vin = nullptr;

    void allocate(std::size_t n_, std::size_t listWords_, int iterations) {
        const std::size_t builds = static_cast<std::size_t>(iterations) + 1;
        if (arena.size() > 0 && n >= n_ && nbr.size() == builds && listWords >= listWords_)
            return;
        // Sizes in 16-byte units, every array starting on a 16-byte boundary.
        const auto units = [](std::size_t bytes) { return (bytes + 15) / 16; };
        const std::size_t u32 = units(n_ * sizeof(std::uint32_t));
        const std::size_t lists = units(listWords_ * sizeof(std::uint32_t));
        const std::size_t total = builds * (2 * u32 + lists) + 2 * (builds - 1) * n_ + 5 * n_;
        arena.allocate(total);
        float4* at = arena.data();
        const auto take = [&](std::size_t count) {
            float4* p = at;
            at += count;
            return p;
        };
        ids.assign(builds, nullptr);
        nbr.assign(builds, nullptr);
        cnt.assign(builds, nullptr);
        pk.assign(builds - 1, nullptr);
        pre.assign(builds - 1, nullptr);
        for (std::size_t b = 0; b < builds; ++b) {
            ids[b] = reinterpret_cast<std::uint32_t*>(take(u32));
            nbr[b] = reinterpret_cast<std::uint32_t*>(take(lists));
            cnt[b] = reinterpret_cast<int*>(take(u32));
        }
        for (std::size_t k = 0; k + 1 < builds; ++k) {
            pk[k] = take(n_);
            pre[k] = take(n_);
        }
        xs = take(n_);
        pf = take(n_);
        v3 = take(n_);
        omega = take(n_);
        vin = take(n_);
        n = n_;
        listWords = listWords_;
    }
};

struct PBFSolver::State {
    std::size_t maxParticles = 0;
    std::size_t count = 0;
    SimulationParams params{};
    GridCore grid;
    CollisionSystem collision;

    // pq holds a skin substep's predicted positions (the lists' reference
    // points) while the iterations ping-pong p and pa.
    DeviceBuffer<float4> x, xa, p, pa, pq, v, va, vin, vina, omega;
    // status[0]: the status word; status[1]: frames finished before a
    // stale one.
    DeviceBuffer<std::uint32_t> id, ida, status;
    // Grid builds inside the constraint loop permute p* alone (lambda and
    // Delta-p read nothing else); x, v, v_in and id follow at the substep's
    // last build.  While they lag, perm[i] is the slot of their records
    // for p* slot i.
    DeviceBuffer<std::uint32_t> perm, perma;
    bool recordsLag = false;
    DeviceBuffer<std::uint32_t> nbr;      // grouped neighbour lists of the last build
    DeviceBuffer<int> cnt;
    std::size_t stride = 0;

    // Skin lists: skin > 0 builds the grid once per substep, on cells of
    // h * (1 + skin) (sgrid); a frame whose lists went stale is run again
    // from the checkpoint with a build per iteration.
    float skin = 0.0f;
    GridCore sgrid;
    DeviceBuffer<pbfint::SceneGeom> sceneGeomSkin;
    DeviceBuffer<float4> ckx, ckv;
    DeviceBuffer<std::uint32_t> ckid;
    bool forceExact = false;
    std::size_t staleFrames = 0;

    // Scenes of a batched solver (PBFBatchSolver); none for one scene.  Each
    // scene keeps its grid, container and slot range in device tables.
    std::size_t sceneCount = 0;
    std::vector<std::size_t> sceneFirst;  // sceneCount + 1 slot offsets
    std::vector<pbf_container> sceneBoxes;
    DeviceBuffer<pbfint::SceneGeom> sceneGeom;
    DeviceBuffer<pbf_container> sceneContainers;
    DeviceBuffer<std::uint32_t> sceneOf;

    pbfint::SceneTable scenes(bool sk = false) const {
        if (sceneCount == 0) return pbfint::SceneTable{};
        return pbfint::SceneTable{sk ? sceneGeomSkin.data() : sceneGeom.data(),
                                  sceneContainers.data(), sceneOf.data()};
    }
    GridCore& grid_of(bool sk) { return sk ? sgrid : grid; }
    // Workgroup size of the neighbour passes (a tuning: results do not
    // depend on it).
    int passBlock = 256;
    int passPpt = 1;  // particles per thread of the neighbour passes (a tuning)
    // Grid builds re-sort from the previous build when the particles are in
    // its order (the same result as a full sort, bit for bit).
    bool resortBuilds = true;
    // XSPH runs inside the post-solve build (the same results bit for bit).
    bool fusedXsph = true;
    // Nothing is known of the order of the particles (new particles, a
    // restored checkpoint): the next builds sort in full.
    void forget_order() {
        grid.orderedCount = 0;
        sgrid.orderedCount = 0;
    }
    pbfint::ListPass lpass() const {
        pbfint::ListPass lp;
        lp.block = passBlock;
        lp.ppt = passPpt;
        return lp;
    }
    bool skin_active() const {
        return skin > 0.0f && tapes == nullptr && !forceExact && !adjointExact;
    }
    bool adjointExact = false;  // the adjoint differentiates the reference dataflow
    // (Re)builds the skin grid for the current skin, bounds and scenes.
    void make_skin_grid(std::size_t capacity);

    void allocate(std::size_t maxParticles_, std::size_t stride_);

    hipStream_t stream = nullptr;
    Profiler prof;
    PinnedStage stage;

    // Frames replayed from HIP graphs.  A frame permutes the working-set
    // buffers (every grid build swaps them with their alternates, the
    // velocity update swaps x and p*), and a graph bakes in the pointers it
    // was captured with, so one graph is kept per starting assignment,
    // with the permutation the frame applies; a workload cycles through one
    // or two of them.  Kernel arguments are baked in too: initialize,
    // setParticles and the collider setters drop the graphs.
    static constexpr int kSlots = 13;
    using BufferKey = std::array<const void*, kSlots>;
    struct FrameGraph {
        BufferKey start{};
        bool skin = false;               // a frame with skin lists
        // Whether the particles start (and end) in the order of the grid's
        // and the skin grid's last sort: the first build re-sorts or not.
        std::array<bool, 2> orderedIn{}, orderedOut{};
        std::array<int, kSlots> perm{};  // slot k ends up holding slot perm[k]'s buffer
        hipGraphExec_t exec = nullptr;
    };
    std::array<bool, 2> ordered() const {
        return {count != 0 && grid.orderedCount == count,
                count != 0 && sgrid.orderedCount == count};
    }
    void set_ordered(const std::array<bool, 2>& o) {
        grid.orderedCount = o[0] ? count : 0;
        sgrid.orderedCount = o[1] ? count : 0;
    }
    std::vector<FrameGraph> graphs;
    bool useGraphs = false;

    // Set while the adjoint replays a checkpoint segment: every substep of
    // the frames records into tapes[tapeNext++].
    std::vector<SubstepTape>* tapes = nullptr;
    std::size_t tapeNext = 0;

    template <typename T>
    void tape_copy(DeviceBuffer<T>& dst, const DeviceBuffer<T>& src, std::size_t count) {
        hip_check(hipMemcpyAsync(dst.data(), src.data(), count * sizeof(T),
                                 hipMemcpyDeviceToDevice, stream));
    }
    SubstepTape* tape() { return tapes != nullptr ? &(*tapes)[tapeNext] : nullptr; }
    void tape_build(int b) {
        SubstepTape* tp = tape();
        if (tp == nullptr) return;
        if (b < 0 || static_cast<std::size_t>(b) >= tp->ids.size())
            throw std::logic_error("Adjoint tape holds fewer builds than the substep runs");
        launch_check(pbfint::tape_build(nbr.data(), tp->nbr[b], cnt.data(), tp->cnt[b], id.data(),
                                        tp->ids[b], count, stride, kMaxNeighbors / 4, stream));
    }

    ~State() {
        drop_graphs();
        if (stream != nullptr) (void)hipStreamDestroy(stream);
    }

    void drop_graphs() noexcept {
        for (FrameGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    // Slots 0-8: x, xa, p, pa, pq, v, va, vin, vina; 9-12: id, ida, perm,
    // perma.
    static constexpr int kF4Slots = 9;
    static constexpr int kU32Slots = kSlots - kF4Slots;
    BufferKey buffer_key() const {
        return {x.data(),  xa.data(),  p.data(),   pa.data(),  pq.data(),
                v.data(),  va.data(),  vin.data(), vina.data(), id.data(),
                ida.data(), perm.data(), perma.data()};
    }

    void permute(const std::array<int, kSlots>& to) {
        std::array<DeviceBuffer<float4>*, kF4Slots> f{&x, &xa, &p, &pa, &pq, &v, &va, &vin, &vina};
        std::array<DeviceBuffer<std::uint32_t>*, kU32Slots> u{&id, &ida, &perm, &perma};
        std::array<DeviceBuffer<float4>, kF4Slots> tf;
        std::array<DeviceBuffer<std::uint32_t>, kU32Slots> tu;
        for (int k = 0; k < kF4Slots; ++k) tf[k] = std::move(*f[to[k]]);
        for (int k = 0; k < kU32Slots; ++k) tu[k] = std::move(*u[to[kF4Slots + k] - kF4Slots]);
        for (int k = 0; k < kF4Slots; ++k) *f[k] = std::move(tf[k]);
        for (int k = 0; k < kU32Slots; ++k) *u[k] = std::move(tu[k]);
    }

    // The permutation that takes assignment `from` to assignment `to`.
    static std::array<int, kSlots> perm_between(const BufferKey& from, const BufferKey& to) {
        std::array<int, kSlots> perm{};
        for (int k = 0; k < kSlots; ++k) {
            const int lo = k < kF4Slots ? 0 : kF4Slots, hi = k < kF4Slots ? kF4Slots : kSlots;
            perm[k] = -1;
            for (int j = lo; j < hi; ++j)
                if (from[j] == to[k]) perm[k] = j;
            if (perm[k] < 0) throw std::logic_error("Working-set buffers lost track of a slot");
        }
        return perm;
    }

    // One frame on the stream: the captured graph of this buffer
    // assignment, captured now if there is none; launched eagerly when
    // graphs are off or the step is being profiled (events per launch).
    void frame() {
        if (!useGraphs || prof.enabled || tapes != nullptr || forceExact) {
            enqueue_frame();
            return;
        }
        const BufferKey key = buffer_key();
        const bool sk = skin_active();
        const std::array<bool, 2> ord = ordered();
        for (const FrameGraph& g : graphs) {
            if (g.start == key && g.skin == sk && g.orderedIn == ord) {
                hip_check(hipGraphLaunch(g.exec, stream));
                permute(g.perm);
                set_ordered(g.orderedOut);
                return;
            }
        }
        FrameGraph fg;
        fg.start = key;
        fg.skin = sk;
        fg.orderedIn = ord;
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_frame();
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph != nullptr) (void)hipGraphDestroy(graph);
            permute(perm_between(buffer_key(), key));
            recordsLag = false;  // nothing ran: every record is where the frame found it
            set_ordered(ord);
            throw;
        }
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        fg.perm = perm_between(key, buffer_key());
        fg.orderedOut = ordered();
        permute(perm_between(buffer_key(), key));  // nothing ran yet
        set_ordered(ord);
        if (ce != hipSuccess) {
            if (graph != nullptr) (void)hipGraphDestroy(graph);
            hip_check(ce);
        }
        const hipError_t ie = hipGraphInstantiate(&fg.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(ie);
        graphs.push_back(fg);
        hip_check(hipGraphLaunch(fg.exec, stream));
        permute(fg.perm);
        set_ordered(fg.orderedOut);
    }

    void ensure_stream() {
        if (stream == nullptr) hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    }

    // Grid build of the current p*, then the permutation of the working set
    // into cell order (SpatialGrid::build + reorderWorkingSet,
    // pbf_solver.cu:154-172, spatial_grid.cu:139-226).  The keys of p* are
    // already in grid.keys and cellStart is already cleared: the prediction
    // kernel does both for a substep's first build, each Delta-p pass for
    // the build after it.
    //
    // The particles sit in the order of this grid's last sort when that
    // sort's build was the last one (orderedCount): the keys are then
    // re-sorted from it, which gives the radix sort's result bit for bit.
    void build_and_reorder(bool sk = false, bool deferRecords = false) {
        const std::size_t n = count;
        GridCore& grid = grid_of(sk);
        grid_of(!sk).orderedCount = 0;  // its order is about to be lost
        {
            const bool re = resortBuilds && grid.orderedCount == n;
            Scope sc(prof, re ? "resort" : "radix_sort", stream);
            grid.sortKeys(n, re, stream);
        }
        {
            Scope sc(prof, "ranges_gather", stream);
            pbfint::Permute pm{grid.sortedIndices.data(), x.data(), p.data(), v.data(),
                               vin.data(), id.data(), xa.data(), pa.data(), va.data(),
                               vina.data(), ida.data()};
            pm.permIn = recordsLag ? perm.data() : nullptr;
            if (deferRecords) {
                pm.x = pm.v = pm.vin = nullptr;
                pm.id = nullptr;
                pm.xo = pm.vo = pm.vino = nullptr;
                pm.ido = nullptr;
                pm.permOut = perma.data();
            }
            launch_check(pbfint::launch_ranges_gather(grid.sortedKeys.data(),
                                                      grid.cellStart.data(), grid.cellEnd.data(),
                                                      n, pm, stream, grid.splitters.data()));
        }
        grid.orderedCount = n;  // p* (and what follows it) is in the new order
        std::swap(p, pa);
        if (deferRecords) {
            std::swap(perm, perma);
            recordsLag = true;
            return;
        }
        std::swap(x, xa);
        std::swap(v, va);
        std::swap(vin, vina);
        std::swap(id, ida);
        recordsLag = false;
    }

    pbfint::SweepGrid sweep_grid(bool sk = false) {
        GridCore& grid = grid_of(sk);
        pbfint::SweepGrid g{grid.cellStart.data(), grid.cellEnd.data(), grid.sortedKeys.data(),
                            grid.dims, count,
                            sk ? grid.cellSize : params.smoothingRadius, kMaxNeighbors,
                            scenes(sk)};
        if (sk) {
            g.skin = true;
            g.box = collision.state().container.data();
            g.radius = params.particleRadius;
        }
        return g;
    }

    // Neighbour lists by cell sweep, then the flat list kernels.
    void build_lists(bool sk = false) {
        Scope sc(prof, "build_lists", stream);
        launch_check(pbfint::sweep_build_lists(sweep_grid(sk), p.data(), nbr.data(), cnt.data(),
                                               stride, status.data(), stream));
    }
    // The post-solve build with XSPH: the lists, and the velocities after
    // applyXsphViscosity into va (and the tape's copy when taping).
    void build_lists_xsph(float visc, float invDt, float4* tapeOut) {
        Scope sc(prof, "build_lists_xsph", stream);
        pbfint::SweepGrid g = sweep_grid();
        g.xs = x.data();
        g.vold = v.data();
        g.vout = va.data();
        g.tapeOut = tapeOut;
        g.visc = visc;
        g.inv_dt = invDt;
        launch_check(pbfint::sweep_build_lists(g, p.data(), nbr.data(), cnt.data(), stride,
                                               status.data(), stream));
    }

    // One frame: params.substeps substeps (pbf_solver.cu:186-325).
    void enqueue_frame() {
        if (skin_active())
            enqueue_frame_skin();
        else
            enqueue_frame_exact();
    }

    // The frame with skin lists: per substep one grid build, around the
    // container points of the predicted positions, with radius
    // h * (1 + skin).  The passes run the reference's sums over the skin
    // lists: every entry is tested against h, and an entry beyond h adds
    // exact zeros (W, grad W and s_corr vanish there), so a pass gives the
    // sums of the exact list in the skin list's order.  This is the
    // reference's result as long as each list holds every pair within h at
    // the positions of each pass, which holds while no particle has moved
    // more than 0.45 * skin * h from its reference point (two such
    // particles within h of each other were within h + 0.9 * skin * h of
    // each other when the list was built).  Each Delta-p pass checks that;
    // a particle past it raises PBF_STATUS_LIST_STALE, the frame stops, and
    // the host runs it again from the checkpoint taken at its start, with a
    // grid build per iteration (finish()).  The counts within h stand in
    // for the counts of the reference's builds: lambda checks those of the
    // builds before each iteration, the first post-solve pass that of the
    // post-solve rebuild.
    void enqueue_frame_skin() {
        const SimulationParams& P = params;
        const std::size_t n = count;
        const float sdt = P.dt / static_cast<float>(P.substeps);
        const float invDt = 1.0f / sdt;
        const float h = P.smoothingRadius;
        const pbfdev::Colliders col = collision.state().view();
        const float half = 0.45f * skin * h;
        recordsLag = false;  // skin builds move every record
        pbfint::ListPass sk = lpass();
        sk.stale2 = half * half;
        sk.exactMax = kMaxNeighbors;
        {
            Scope sc(prof, "checkpoint", stream);
            launch_check(pbfint::launch_checkpoint(x.data(), v.data(), id.data(), ckx.data(),
                                                   ckv.data(), ckid.data(), n, status.data(),
                                                   stream));
        }
        for (int substep = 0; substep < P.substeps; ++substep) {
            {
                Scope sc(prof, "predict_key", stream);
                launch_check(pbfint::launch_predict_key(
                    x.data(), p.data(), v.data(), vin.data(), sgrid.keys.data(),
                    static_cast<int>(n), sdt, P.gravity, sgrid.minBounds, sgrid.cellSize,
                    sgrid.dims, col.container, P.particleRadius, scenes(true),
                    sgrid.cellStart.data(), sgrid.numCells, status.data(), stream));
            }
            build_and_reorder(true);
            build_lists(true);
            for (int it = 0; it < P.solverIterations; ++it) {
                sk.p0 = it == 0 ? p.data() : pq.data();
                {
                    Scope sc(prof, "lambda", stream);
                    launch_check(pbfint::launch_lambda_guarded(
                        p.data(), nbr.data(), cnt.data(), kMaxNeighbors, n, h, P.particleMass,
                        P.restDensity, P.lambdaEpsilon, stride, sk, status.data(), stream));
                }
                {
                    Scope sc(prof, "delta_apply_collide", stream);
                    launch_check(pbfint::launch_delta_apply_collide(
                        p.data(), pa.data(), nbr.data(), cnt.data(), kMaxNeighbors, x.data(),
                        n, h, P.particleMass, P.restDensity, P.scorrK, P.scorrN, P.scorrDeltaQ,
                        stride, col, P.particleRadius, scenes(), sk, status.data(), stream));
                }
                std::swap(p, pa);
                if (it == 0) std::swap(pa, pq);  // pq keeps the predicted positions
            }
            const bool post = P.xsphViscosity != 0.0f || P.vorticityStrength != 0.0f;
            pbfint::ListPass postSkin = lpass();
            postSkin.exactMax = post ? kMaxNeighbors : 0;
            if (P.xsphViscosity != 0.0f) {
                Scope sc(prof, "update_xsph", stream);
                launch_check(pbfint::launch_update_xsph(
                    x.data(), p.data(), v.data(), va.data(), nbr.data(), cnt.data(), kMaxNeighbors,
                    n, h, P.xsphViscosity, invDt, stride, postSkin, status.data(), stream));
                std::swap(v, va);
                postSkin.exactMax = 0;
            } else {
                Scope sc(prof, "update_velocity", stream);
                launch_check(pbfint::launch_update_velocity_only(x.data(), p.data(), v.data(), n,
                                                                 invDt, status.data(), stream));
            }
            std::swap(x, p);
            if (P.vorticityStrength != 0.0f) {
                {
                    Scope sc(prof, "vorticity", stream);
                    launch_check(pbfint::launch_vorticity(
                        x.data(), v.data(), nbr.data(), cnt.data(), kMaxNeighbors, n, h,
                        omega.data(), p.data(), stride, postSkin, status.data(), stream));
                }
                {
                    Scope sc(prof, "confine_resolve", stream);
                    launch_check(pbfint::launch_confine_resolve(
                        x.data(), p.data(), nbr.data(), cnt.data(), kMaxNeighbors, omega.data(),
                        v.data(), vin.data(), va.data(), n, h, sdt, P.vorticityStrength, stride, col,
                        P.particleRadius, P.collisionRestitution, P.collisionFriction, scenes(),
                        lpass(), status.data(), stream));
                }
                std::swap(v, va);
            } else {
                Scope sc(prof, "resolve_velocities", stream);
                launch_check(pbfint::launch_resolve_guarded(
                    x.data(), vin.data(), v.data(), n, col, P.particleRadius,
                    P.collisionRestitution, P.collisionFriction, scenes(), status.data(), stream));
            }
        }
        launch_check(pbfint::launch_frame_done(status.data(), stream));
    }

    // One frame with the reference's dataflow: a grid build before every
    // constraint iteration and after the solve.
    void enqueue_frame_exact() {
        const SimulationParams& P = params;
        const std::size_t n = count;
        const float sdt = P.dt / static_cast<float>(P.substeps);
        const float invDt = 1.0f / sdt;
        const float h = P.smoothingRadius;
        const pbfdev::Colliders col = collision.state().view();
        // The reference rebuilds after the solve only for XSPH or vorticity
        // (pbf_solver.cu:260).  Under the adjoint every frame rebuilds, taped
        // or not, so that a replayed segment follows the frames forward()
        // ran, and the derivatives with respect to c and epsilon at 0 have
        // the final positions' lists.
        const bool post =
            P.xsphViscosity != 0.0f || P.vorticityStrength != 0.0f || adjointExact;
        recordsLag = false;
        for (int substep = 0; substep < P.substeps; ++substep) {
            {
                Scope sc(prof, "predict_key", stream);
                launch_check(pbfint::launch_predict_key(
                    x.data(), p.data(), v.data(), vin.data(), grid.keys.data(),
                    static_cast<int>(n), sdt, P.gravity, grid.minBounds, grid.cellSize, grid.dims,
                    nullptr, 0.0f, scenes(), grid.cellStart.data(), grid.numCells, status.data(),
                    stream));
            }
            for (int it = 0; it < P.solverIterations; ++it) {
                // The substep's last build moves every record; the tapes
                // record ids at every build.
                const bool last = !post && it == P.solverIterations - 1;
                build_and_reorder(false, !last && tapes == nullptr);
                build_lists();
                tape_build(it);
                {
                    Scope sc(prof, "lambda", stream);
                    pbfint::ListPass ll = lpass();
                    if (SubstepTape* tp = tape()) ll.tapeOut = tp->pk[it];
                    launch_check(pbfint::launch_lambda_guarded(
                        p.data(), nbr.data(), cnt.data(), kMaxNeighbors, n, h, P.particleMass,
                        P.restDensity, P.lambdaEpsilon, stride, ll, status.data(), stream));
                }
                {
                    Scope sc(prof, "delta_apply_collide", stream);
                    pbfint::ListPass dl = lpass();
                    dl.keyOut = grid.keys.data();  // the next build's keys and cells
                    dl.keyMin = grid.minBounds;
                    dl.keyCell = grid.cellSize;
                    dl.keyDims = grid.dims;
                    dl.clearCells = grid.cellStart.data();
                    dl.clearCount = grid.numCells;
                    dl.xperm = recordsLag ? perm.data() : nullptr;
                    if (SubstepTape* tp = tape()) dl.prePos = tp->pre[it];
                    launch_check(pbfint::launch_delta_apply_collide(
                        p.data(), pa.data(), nbr.data(), cnt.data(), kMaxNeighbors, x.data(),
                        n, h, P.particleMass, P.restDensity, P.scorrK, P.scorrN, P.scorrDeltaQ,
                        stride, col, P.particleRadius, scenes(), dl, status.data(), stream));
                }
                std::swap(p, pa);
            }
            const bool fuseXsph = fusedXsph && P.xsphViscosity != 0.0f && post;
            if (post) {
                build_and_reorder();
                if (fuseXsph) {
                    SubstepTape* tp = tape();
                    build_lists_xsph(P.xsphViscosity, invDt, tp != nullptr ? tp->v3 : nullptr);
                } else {
                    build_lists();
                }
                tape_build(P.solverIterations);
                if (SubstepTape* tp = tape())
                    launch_check(pbfint::tape_records(x.data(), tp->xs, p.data(), tp->pf,
                                                      vin.data(), tp->vin, n, stream));
            }
            if (fuseXsph) {
                std::swap(v, va);  // the build wrote the velocities
            } else if (P.xsphViscosity != 0.0f) {
                Scope sc(prof, "update_xsph", stream);
                pbfint::ListPass xl = lpass();
                if (SubstepTape* tp = tape()) xl.tapeOut = tp->v3;
                launch_check(pbfint::launch_update_xsph(
                    x.data(), p.data(), v.data(), va.data(), nbr.data(), cnt.data(), kMaxNeighbors,
                    n, h, P.xsphViscosity, invDt, stride, xl, status.data(), stream));
                std::swap(v, va);
            } else {
                Scope sc(prof, "update_velocity", stream);
                launch_check(pbfint::launch_update_velocity_only(x.data(), p.data(), v.data(), n,
                                                                 invDt, status.data(), stream));
                if (SubstepTape* tp = tape())
                    hip_check(hipMemcpyAsync(tp->v3, v.data(), n * sizeof(float4),
                                             hipMemcpyDeviceToDevice, stream));
            }
            std::swap(x, p);
            if (P.vorticityStrength != 0.0f) {
                {
                    Scope sc(prof, "vorticity", stream);
                    // p holds the previous x after the swap and is free until
                    // the next substep's prediction: it takes the records.
                    pbfint::ListPass vl = lpass();
                    if (SubstepTape* tp = tape()) vl.tapeOut = tp->omega;
                    launch_check(pbfint::launch_vorticity(
                        x.data(), v.data(), nbr.data(), cnt.data(), kMaxNeighbors, n, h,
                        omega.data(), p.data(), stride, vl, status.data(), stream));
                }
                {
                    Scope sc(prof, "confine_resolve", stream);
                    launch_check(pbfint::launch_confine_resolve(
                        x.data(), p.data(), nbr.data(), cnt.data(), kMaxNeighbors, omega.data(),
                        v.data(), vin.data(), va.data(), n, h, sdt, P.vorticityStrength, stride, col,
                        P.particleRadius, P.collisionRestitution, P.collisionFriction, scenes(),
                        lpass(), status.data(), stream));
                }
                std::swap(v, va);
            } else {
                if (SubstepTape* tp = tape()) {
                    // Confinement is off, but its strength's derivative at 0
                    // needs the vorticity: computed for the tape only (the
                    // pass changes no state; p takes its records).
                    pbfint::ListPass vl = lpass();
                    vl.tapeOut = tp->omega;
                    launch_check(pbfint::launch_vorticity(
                        x.data(), v.data(), nbr.data(), cnt.data(), kMaxNeighbors, n, h,
                        omega.data(), p.data(), stride, vl, status.data(), stream));
                }
                Scope sc(prof, "resolve_velocities", stream);
                launch_check(pbfint::launch_resolve_guarded(
                    x.data(), vin.data(), v.data(), n, col, P.particleRadius,
                    P.collisionRestitution, P.collisionFriction, scenes(), status.data(), stream));
            }
            if (tapes != nullptr) ++tapeNext;
        }
    }

    void reset_status() { launch_check(pbfint::launch_zero_u32(status.data(), 2, stream)); }

    // Host -> device copy of n records, each record checked first: one DMA
    // from pinned sources, else through the pinned stage (the check and the
    // staging copy are one pass over the source, overlapping the previous
    // chunk's DMA).  Throws, with the stream drained, at the first
    // non-finite coordinate.
    void upload_checked(const float4* src, float4* dst, std::size_t n) {
        const HostArray kind = host_array_kind(src);
        if (kind == HostArray::Device)
            throw std::invalid_argument(
                "Particle arrays must be host memory (use setParticlesDevice for device arrays)");
        if (kind == HostArray::Pinned) {
            for (std::size_t k = 0; k < n; ++k)
                if (!finite4(src[k]))
                    throw std::invalid_argument(
                        "Particle coordinates and velocities must be finite");
            hip_check(hipMemcpyAsync(dst, src, n * sizeof(float4), hipMemcpyHostToDevice, stream));
            return;
        }
        const std::size_t per = PinnedStage::kChunk;
        int slot = 0;
        try {
            for (std::size_t i0 = 0; i0 < n; i0 += per, slot ^= 1) {
                const std::size_t m = std::min(per, n - i0);
                hip_check(hipEventSynchronize(stage.event(slot)));
                float4* h = stage.buffer(slot);
                for (std::size_t k = 0; k < m; ++k) {
                    const float4 v = src[i0 + k];
                    if (!finite4(v))
                        throw std::invalid_argument(
                            "Particle coordinates and velocities must be finite");
                    h[k] = v;
                }
                hip_check(hipMemcpyAsync(dst + i0, h, m * sizeof(float4), hipMemcpyHostToDevice,
                                         stream));
                hip_check(hipEventRecord(stage.event(slot), stream));
            }
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            throw;
        }
    }

    // Device -> host copy of n records: straight into pinned destinations,
    // else through the pinned stage, the host copy of one chunk overlapping
    // the DMA of the next.  Returns with the copy complete.
    void download(const float4* src, float4* out, std::size_t n) {
        const HostArray kind = host_array_kind(out);
        if (kind == HostArray::Device)
            throw std::invalid_argument(
                "Destination must be host memory (use the copy-to-device calls for device arrays)");
        if (kind == HostArray::Pinned) {
            hip_check(hipMemcpyAsync(out, src, n * sizeof(float4), hipMemcpyDeviceToHost, stream));
            return;
        }
        const std::size_t per = PinnedStage::kChunk;
        const std::size_t chunks = (n + per - 1) / per;
        for (std::size_t c = 0; c <= chunks; ++c) {
            if (c < chunks) {
                const std::size_t i0 = c * per;
                const std::size_t m = std::min(per, n - i0);
                const int slot = static_cast<int>(c & 1);
                hip_check(hipMemcpyAsync(stage.buffer(slot), src + i0, m * sizeof(float4),
                                         hipMemcpyDeviceToHost, stream));
                hip_check(hipEventRecord(stage.event(slot), stream));
            }
            if (c > 0) {
                const std::size_t i0 = (c - 1) * per;
                const std::size_t m = std::min(per, n - i0);
                const int slot = static_cast<int>((c - 1) & 1);
                hip_check(hipEventSynchronize(stage.event(slot)));
                std::memcpy(out + i0, stage.buffer(slot), m * sizeof(float4));
            }
        }
    }

    // Waits for the device and raises what the status word recorded.
    // Returns the status word.
    std::uint32_t finish() {
        hip_check(hipStreamSynchronize(stream));
        prof.collect();
        std::uint32_t st[2] = {0, 0};
        hip_check(hipMemcpy(st, status.data(), sizeof(st), hipMemcpyDeviceToHost));
        if (st[0] & PBF_STATUS_NEIGHBOR_OVERFLOW)
            throw std::overflow_error("Particle neighbor count exceeds solver capacity");
        if (st[0] & PBF_STATUS_COLLISION_NONCONVERGENCE)
            throw collision_nonconvergence("Collision constraints did not converge");
        framesDone = st[1];
        return st[0];
    }
    std::uint32_t framesDone = 0;

    // frameCount frames, each a step(): a frame whose skin lists went stale
    // is run again from its checkpoint with a build per iteration, then the
    // rest follow.  Overflow and non-convergence throw as in the reference.
    void run_frames(std::size_t frameCount) {
        while (frameCount > 0) {
            reset_status();
            for (std::size_t f = 0; f < frameCount; ++f) frame();
            if ((finish() & PBF_STATUS_LIST_STALE) == 0u) return;
            const std::size_t done = framesDone;
            const std::size_t n = count;
            hip_check(hipMemcpyAsync(x.data(), ckx.data(), n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, stream));
            hip_check(hipMemcpyAsync(v.data(), ckv.data(), n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, stream));
            hip_check(hipMemcpyAsync(id.data(), ckid.data(), n * sizeof(std::uint32_t),
                                     hipMemcpyDeviceToDevice, stream));
            forget_order();
            ++staleFrames;
            reset_status();
            forceExact = true;
            try {
                frame();
            } catch (...) {
                forceExact = false;
                throw;
            }
            forceExact = false;
            finish();
            frameCount -= done + 1;
        }
    }

    void copy_out(const DeviceBuffer<float4>& src, DeviceBuffer<float4>& scratch, float4* out,
                  std::size_t n, bool toHost) {
        launch_check(pbf_k_scatter_float4_by_stable_id(src.data(), id.data(), scratch.data(), n,
                                                       stream));
        if (toHost) {
            if (out == nullptr) {
                hip_check(hipStreamSynchronize(stream));
                throw std::invalid_argument("Destination cannot be null");
            }
            if (host_array_kind(out) == HostArray::Device) {
                hip_check(hipStreamSynchronize(stream));
                throw std::invalid_argument(
                    "Destination must be host memory (use the copy-to-device calls for device "
                    "arrays)");
            }
            download(scratch.data(), out, n);
        } else {
            if (out == nullptr) {
                hip_check(hipStreamSynchronize(stream));
                throw std::invalid_argument("Destination cannot be null");
            }
            hip_check(hipMemcpyAsync(out, scratch.data(), n * sizeof(float4),
                                     hipMemcpyDeviceToDevice, stream));
        }
        hip_check(hipStreamSynchronize(stream));
    }
};

PBFSolver::PBFSolver() : state_(new State()) {}
PBFSolver::~PBFSolver() = default;
PBFSolver::PBFSolver(PBFSolver&&) noexcept = default;
PBFSolver& PBFSolver::operator=(PBFSolver&&) noexcept = default;

namespace {

// Capacity checks of pbf_solver.cu:25-43; returns the list stride.
std::size_t checked_stride(std::size_t maxParticles) {
    if (maxParticles == 0) throw std::invalid_argument("maxParticles must be greater than zero");
    if (maxParticles > std::numeric_limits<std::size_t>::max() / static_cast<std::size_t>(kMaxNeighbors))
        throw std::length_error("Neighbor buffer size overflows size_t");
    if (maxParticles > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::length_error("Particle count exceeds kernel index range");
    const std::size_t stride =
        ((maxParticles + kBlockSize - 1) / kBlockSize) * static_cast<std::size_t>(kBlockSize);
    if (stride > std::numeric_limits<std::size_t>::max() / static_cast<std::size_t>(kMaxNeighbors))
        throw std::length_error("Neighbor buffer size overflows size_t");
    return stride;
}

// Parameter checks of pbf_solver.cu:45-79.
void check_params(const SimulationParams& p) {
    if (!std::isfinite(p.dt) || p.dt <= 0.0f || !std::isfinite(p.restDensity) ||
        p.restDensity <= 0.0f || !std::isfinite(p.particleMass) || p.particleMass <= 0.0f ||
        !std::isfinite(p.particleRadius) || p.particleRadius <= 0.0f ||
        !std::isfinite(p.collisionRestitution) || p.collisionRestitution < 0.0f ||
        p.collisionRestitution > 1.0f || !std::isfinite(p.collisionFriction) ||
        p.collisionFriction < 0.0f || p.collisionFriction > 1.0f ||
        !std::isfinite(p.smoothingRadius) || p.smoothingRadius <= 0.0f ||
        !std::isfinite(p.lambdaEpsilon) || p.lambdaEpsilon <= 0.0f || !std::isfinite(p.scorrK) ||
        p.scorrK < 0.0f ||
        (p.scorrK > 0.0f && (p.scorrN <= 0 || !std::isfinite(p.scorrDeltaQ) ||
                             p.scorrDeltaQ < 0.0f || p.scorrDeltaQ >= p.smoothingRadius)) ||
        !std::isfinite(p.xsphViscosity) || p.xsphViscosity < 0.0f ||
        !std::isfinite(p.vorticityStrength) || p.vorticityStrength < 0.0f ||
        !std::isfinite(p.gravity.x) || !std::isfinite(p.gravity.y) || !std::isfinite(p.gravity.z) ||
        p.solverIterations <= 0 || p.substeps <= 0) {
        throw std::invalid_argument("Invalid PBF simulation parameters");
    }
    const float substepDt = p.dt / static_cast<float>(p.substeps);
    if (!std::isfinite(substepDt) || substepDt <= 0.0f || !std::isfinite(1.0f / substepDt))
        throw std::invalid_argument("Substep duration is not representable");
}

// The closed container must fit a particle (pbf_solver.cu:81-82).
void check_container(float3 minBounds, float3 maxBounds, const SimulationParams& p) {
    const double width = static_cast<double>(maxBounds.x) - static_cast<double>(minBounds.x);
    const double depth = static_cast<double>(maxBounds.z) - static_cast<double>(minBounds.z);
    const double diameter = 2.0 * static_cast<double>(p.particleRadius);
    if (width < diameter || depth < diameter)
        throw std::invalid_argument("Particle diameter exceeds closed container dimensions");
}

// The grids of a batch's scenes (their bounds, cells of size cs) laid end
// to end: global cell offsets and slot ranges.
std::vector<pbfint::SceneGeom> scene_grids(const std::vector<pbf_container>& boxes,
                                           const std::vector<std::size_t>& first, float cs,
                                           std::size_t* cellsOut) {
    std::vector<pbfint::SceneGeom> geom(boxes.size());
    std::size_t cells = 0;
    for (std::size_t k = 0; k < boxes.size(); ++k) {
        std::size_t sceneCells = 0;
        const int3 d = GridCore::dimsFor(boxes[k].min, boxes[k].max, cs, &sceneCells);
        if (cells + sceneCells > std::numeric_limits<std::uint32_t>::max())
            throw std::length_error("Batched grids exceed 32-bit key capacity");
        pbfint::SceneGeom& g = geom[k];
        g.minBounds = boxes[k].min;
        g.dims = d;
        g.cellOff = static_cast<std::uint32_t>(cells);
        g.first = static_cast<std::uint32_t>(first[k]);
        g.end = static_cast<std::uint32_t>(first[k + 1]);
        cells += sceneCells;
    }
    *cellsOut = cells;
    return geom;
}

}  // namespace

// The grid of the skin lists: the scene's (or every scene's) bounds on
// cells of h * (1 + skin), the list radius, so the lists keep the 3x3x3
// stencil.  Built whole before it replaces the old one.
void PBFSolver::State::make_skin_grid(std::size_t capacity) {
    GridCore next;
    DeviceBuffer<pbfint::SceneGeom> nextGeom;
    if (skin > 0.0f && capacity > 0) {
        const float cs = params.smoothingRadius * (1.0f + skin);
        if (sceneCount == 0) {
            next = GridCore::create(capacity, grid.minBounds, grid.maxBounds, cs);
        } else {
            std::size_t cells = 0;
            const std::vector<pbfint::SceneGeom> geom =
                scene_grids(sceneBoxes, sceneFirst, cs, &cells);
            next = GridCore::createKeyed(capacity, cells, cs);
            nextGeom.allocate(geom.size());
            nextGeom.copyFromHostToDevice(geom.data(), geom.size());
        }
    }
    drop_graphs();
    sgrid = std::move(next);
    sceneGeomSkin = std::move(nextGeom);
}

// Allocates the working set and lists for maxParticles particles (after
// the grid is in place); the solver stays inactive until the end.
void PBFSolver::State::allocate(std::size_t maxParticles_, std::size_t stride_) {
    for (DeviceBuffer<float4>* b : {&x, &xa, &p, &pa, &pq, &v, &va, &vin, &vina, &omega, &ckx, &ckv})
        b->allocate(maxParticles_);
    id.allocate(maxParticles_);
    ida.allocate(maxParticles_);
    perm.allocate(maxParticles_);
    perma.allocate(maxParticles_);
    ckid.allocate(maxParticles_);
    status.allocate(2);
    nbr.allocate(stride_ * static_cast<std::size_t>(kMaxNeighbors));
    cnt.allocate(maxParticles_);
    stride = stride_;
}

void PBFSolver::initialize(std::size_t maxParticles, float3 minBounds, float3 maxBounds,
                           const SimulationParams& params) {
    // Validation order and exception types of pbf_solver.cu:25-79.  The
    // neighbour-list bounds are checked like the reference's, so the same
    // inputs fail the same way.
    const std::size_t stride = checked_stride(maxParticles);
    check_params(params);
    check_container(minBounds, maxBounds, params);

    State& s = *state_;
    GridCore grid = GridCore::create(maxParticles, minBounds, maxBounds, params.smoothingRadius);
    s.drop_graphs();

    // From here a failed allocation leaves the solver inactive
    // (pbf_solver.cu:84-87).
    s.maxParticles = 0;
    s.count = 0;
    s.sceneCount = 0;
    s.sceneFirst.clear();
    s.sceneGeom = DeviceBuffer<pbfint::SceneGeom>();
    s.sceneContainers = DeviceBuffer<pbf_container>();
    s.sceneOf = DeviceBuffer<std::uint32_t>();
    s.grid = std::move(grid);
    s.collision.setContainer(Container{minBounds, maxBounds});
    s.ensure_stream();
    s.allocate(maxParticles, stride);
    s.params = params;
    s.staleFrames = 0;
    s.make_skin_grid(maxParticles);
    s.maxParticles = maxParticles;
    s.count = 0;
}

void PBFSolver::setParticles(const float4* positions, const float4* velocities, std::size_t n) {
    State& s = *state_;
    if (s.maxParticles == 0)
        throw std::logic_error("PBFSolver must be initialized before setting particles");
    if (n > s.maxParticles) throw std::out_of_range("Particle count exceeds solver capacity");
    if (n == 0) {
        s.count = 0;
        return;
    }
    if (positions == nullptr || velocities == nullptr)
        throw std::invalid_argument("Particle positions and velocities must not be null");
    s.drop_graphs();
    // Validated while staged into the alternate buffers, which hold nothing
    // between steps: a non-finite value throws with x and v untouched
    // (pbf_solver.cu:131-143 validates both arrays before uploading).
    s.upload_checked(positions, s.xa.data(), n);
    s.upload_checked(velocities, s.va.data(), n);
    launch_check(pbf_k_initialize_stable_particle_ids(s.id.data(), n, s.stream));
    hip_check(hipStreamSynchronize(s.stream));
    std::swap(s.x, s.xa);
    std::swap(s.v, s.va);
    s.count = n;
    s.forget_order();
}

void PBFSolver::setParticlesDevice(const float4* positions, const float4* velocities,
                                   std::size_t n) {
    State& s = *state_;
    if (s.maxParticles == 0)
        throw std::logic_error("PBFSolver must be initialized before setting particles");
    if (n > s.maxParticles) throw std::out_of_range("Particle count exceeds solver capacity");
    if (n == 0) {
        s.count = 0;
        return;
    }
    if (positions == nullptr || velocities == nullptr)
        throw std::invalid_argument("Particle positions and velocities must not be null");
    s.drop_graphs();
    hip_check(hipDeviceSynchronize());
    hip_check(hipMemcpyAsync(s.x.data(), positions, n * sizeof(float4), hipMemcpyDeviceToDevice,
                             s.stream));
    hip_check(hipMemcpyAsync(s.v.data(), velocities, n * sizeof(float4), hipMemcpyDeviceToDevice,
                             s.stream));
    launch_check(pbf_k_initialize_stable_particle_ids(s.id.data(), n, s.stream));
    hip_check(hipStreamSynchronize(s.stream));
    s.count = n;
    s.forget_order();
}

void PBFSolver::setSpheres(const std::vector<SphereCollider>& spheres) {
    state_->collision.setSpheres(spheres);
    state_->drop_graphs();
}
void PBFSolver::setBoxes(const std::vector<BoxCollider>& boxes) {
    state_->collision.setBoxes(boxes);
    state_->drop_graphs();
}
void PBFSolver::setPlanes(const std::vector<PlaneCollider>& planes) {
    state_->collision.setPlanes(planes);
    state_->drop_graphs();
}

void PBFSolver::step() {
    State& s = *state_;
    if (s.maxParticles == 0) throw std::logic_error("PBFSolver must be initialized before stepping");
    if (s.count == 0) return;
    s.run_frames(1);
}

void PBFSolver::run() {
    if (state_->maxParticles == 0)
        throw std::logic_error("PBFSolver must be initialized before running");
    step();
}

void PBFSolver::run(std::size_t frameCount) {
    State& s = *state_;
    if (s.maxParticles == 0) throw std::logic_error("PBFSolver must be initialized before running");
    if (s.count == 0 || frameCount == 0) return;
    s.run_frames(frameCount);
}

void PBFSolver::copyPositionsToHost(float4* out, std::size_t n) const {
    State& s = *state_;
    if (n != s.count)
        throw std::out_of_range("Requested position count does not match active particles");
    if (n != 0) s.copy_out(s.x, s.xa, out, n, true);
}

void PBFSolver::copyVelocitiesToHost(float4* out, std::size_t n) const {
    State& s = *state_;
    if (n != s.count)
        throw std::out_of_range("Requested velocity count does not match active particles");
    if (n != 0) s.copy_out(s.v, s.va, out, n, true);
}

void PBFSolver::copyPositionsToDevice(float4* out, std::size_t n) const {
    State& s = *state_;
    if (n != s.count)
        throw std::out_of_range("Requested position count does not match active particles");
    if (n != 0) s.copy_out(s.x, s.xa, out, n, false);
}

void PBFSolver::copyVelocitiesToDevice(float4* out, std::size_t n) const {
    State& s = *state_;
    if (n != s.count)
        throw std::out_of_range("Requested velocity count does not match active particles");
    if (n != 0) s.copy_out(s.v, s.va, out, n, false);
}

std::size_t PBFSolver::particleCount() const { return state_->count; }
std::size_t PBFSolver::maxParticles() const { return state_->maxParticles; }
void* PBFSolver::stream() const { return state_->stream; }

void PBFSolver::setNeighborSkin(float skin) {
    if (!std::isfinite(skin) || skin < 0.0f || skin > 1.0f)
        throw std::invalid_argument("Neighbour skin must lie in [0, 1]");
    State& s = *state_;
    const float old = s.skin;
    s.skin = skin;
    try {
        s.make_skin_grid(s.maxParticles);
    } catch (...) {
        s.skin = old;
        throw;
    }
}

float PBFSolver::neighborSkin() const { return state_->skin; }
std::size_t PBFSolver::staleFrames() const { return state_->staleFrames; }

std::vector<PBFSolver::Tuning> PBFSolver::defaultTunings() {
    std::vector<Tuning> out;
    for (float skin : {0.0f, 0.1f, 0.2f, 0.3f})
        for (int block : {64, 128, 256})
            for (int ppt : {1, 2}) out.push_back(Tuning{skin, block, ppt});
    return out;
}

namespace {
void check_tuning(const PBFSolver::Tuning& t) {
    if (!std::isfinite(t.neighborSkin) || t.neighborSkin < 0.0f || t.neighborSkin > 1.0f)
        throw std::invalid_argument("Neighbour skin must lie in [0, 1]");
    if (t.passBlock < 64 || t.passBlock > 256 || t.passBlock % 64 != 0)
        throw std::invalid_argument("passBlock must be 64, 128, 192 or 256");
    if (t.particlesPerThread != 1 && t.particlesPerThread != 2 && t.particlesPerThread != 4)
        throw std::invalid_argument("particlesPerThread must be 1, 2 or 4");
}
}  // namespace

void PBFSolver::setTuning(const Tuning& tuning) {
    check_tuning(tuning);
    setNeighborSkin(tuning.neighborSkin);
    state_->passBlock = tuning.passBlock;
    state_->passPpt = tuning.particlesPerThread;
    state_->drop_graphs();
}

PBFSolver::Tuning PBFSolver::tuning() const {
    return Tuning{state_->skin, state_->passBlock, state_->passPpt};
}

std::vector<PBFSolver::TuningTime> PBFSolver::autotune(std::size_t frames,
                                                       const std::vector<Tuning>& candidates) {
    State& s = *state_;
    if (s.maxParticles == 0)
        throw std::logic_error("PBFSolver must be initialized before autotuning");
    if (frames == 0) throw std::invalid_argument("The autotuner needs at least one frame");
    const std::vector<Tuning> cands = candidates.empty() ? defaultTunings() : candidates;
    for (const Tuning& t : cands) check_tuning(t);
    std::vector<TuningTime> times;
    if (s.count == 0) return times;
    const Tuning original = tuning();
    const std::size_t stale0 = s.staleFrames;
    const std::size_t n = s.count;
    DeviceBuffer<float4> sx, sv;
    DeviceBuffer<std::uint32_t> sid;
    sx.allocate(n);
    sv.allocate(n);
    sid.allocate(n);
    const auto copy = [&s, n](void* dst, const void* src, std::size_t bytes) {
        hip_check(hipMemcpyAsync(dst, src, n * bytes, hipMemcpyDeviceToDevice, s.stream));
    };
    copy(sx.data(), s.x.data(), sizeof(float4));
    copy(sv.data(), s.v.data(), sizeof(float4));
    copy(sid.data(), s.id.data(), sizeof(std::uint32_t));
    const auto restore = [&] {
        copy(s.x.data(), sx.data(), sizeof(float4));
        copy(s.v.data(), sv.data(), sizeof(float4));
        copy(s.id.data(), sid.data(), sizeof(std::uint32_t));
        s.forget_order();
        hip_check(hipStreamSynchronize(s.stream));
    };
    try {
        for (const Tuning& t : cands) {
            setTuning(t);
            restore();
            s.run_frames(1);  // first use: graph capture, caches
            restore();
            const auto t0 = std::chrono::steady_clock::now();
            s.run_frames(frames);
            const auto t1 = std::chrono::steady_clock::now();
            times.push_back(TuningTime{
                t, std::chrono::duration<double, std::milli>(t1 - t0).count() /
                       static_cast<double>(frames)});
        }
    } catch (...) {
        restore();
        setTuning(original);
        s.staleFrames = stale0;
        throw;
    }
    const TuningTime* best = &times.front();
    for (const TuningTime& tt : times)
        if (tt.msPerFrame < best->msPerFrame) best = &tt;
    restore();
    setTuning(best->tuning);
    s.staleFrames = stale0;
    return times;
}

void PBFSolver::setGraphs(bool enabled) {
    state_->useGraphs = enabled;
    if (!enabled) state_->drop_graphs();
}

void PBFSolver::setIncrementalSort(bool enabled) {
    state_->resortBuilds = enabled;
    state_->drop_graphs();
}

void PBFSolver::setFusedXsph(bool enabled) {
    state_->fusedXsph = enabled;
    state_->drop_graphs();
}

void PBFSolver::setProfiling(bool enabled) {
    state_->prof.enabled = enabled;
    state_->prof.reset();
}

std::vector<PBFSolver::KernelTime> PBFSolver::kernelTimes() const { return state_->prof.times(); }

void PBFSolver::copyNeighborLists(float4* positions, int* counts, std::uint32_t* neighbors,
                                  std::size_t n, int maxNeighbors) const {
    State& s = *state_;
    if (s.maxParticles == 0)
        throw std::logic_error("PBFSolver must be initialized before reading neighbour lists");
    if (n != s.count)
        throw std::out_of_range("Requested neighbour list count does not match active particles");
    if (maxNeighbors < 0 || maxNeighbors > kMaxNeighbors)
        throw std::invalid_argument("maxNeighbors must lie in [0, 256]");
    if (n == 0) return;
    if (positions == nullptr || counts == nullptr || (maxNeighbors > 0 && neighbors == nullptr))
        throw std::invalid_argument("Destination cannot be null");
    hip_check(hipStreamSynchronize(s.stream));
    hip_check(hipMemcpy(positions, s.x.data(), n * sizeof(float4), hipMemcpyDeviceToHost));
    hip_check(hipMemcpy(counts, s.cnt.data(), n * sizeof(int), hipMemcpyDeviceToHost));
    if (maxNeighbors > 0) {
        // The step keeps its lists in groups of four (entry q of particle i
        // at word 4 * ((q / 4) * stride + i) + q % 4); hand them out
        // slot-major, as findNeighbors writes them.
        const std::size_t groups = (static_cast<std::size_t>(maxNeighbors) + 3) / 4;
        std::vector<std::uint32_t> g(groups * n * 4);
        hip_check(hipMemcpy2D(g.data(), n * 4 * sizeof(std::uint32_t), s.nbr.data(),
                              s.stride * 4 * sizeof(std::uint32_t), n * 4 * sizeof(std::uint32_t),
                              groups, hipMemcpyDeviceToHost));
        for (std::size_t q = 0; q < static_cast<std::size_t>(maxNeighbors); ++q)
            for (std::size_t i = 0; i < n; ++i)
                neighbors[q * n + i] = g[((q / 4) * n + i) * 4 + q % 4];
    }
}

// ---------------------------------------------------------------------------
// PBFBatchSolver: the same step over many scenes (see pbf/pbf_batch.hpp)
// ---------------------------------------------------------------------------

PBFBatchSolver::PBFBatchSolver() = default;
PBFBatchSolver::~PBFBatchSolver() = default;
PBFBatchSolver::PBFBatchSolver(PBFBatchSolver&&) noexcept = default;
PBFBatchSolver& PBFBatchSolver::operator=(PBFBatchSolver&&) noexcept = default;

void PBFBatchSolver::initialize(const std::vector<BatchScene>& scenes,
                                const SimulationParams& params) {
    if (scenes.empty()) throw std::invalid_argument("A batch needs at least one scene");
    // Capacity of every scene, then of the whole batch (one working set).
    std::size_t total = 0;
    for (const BatchScene& sc : scenes) {
        (void)checked_stride(sc.particleCount);
        if (sc.particleCount > std::numeric_limits<std::size_t>::max() - total)
            throw std::length_error("Particle count exceeds kernel index range");
        total += sc.particleCount;
    }
    const std::size_t stride = checked_stride(total);
    if (total > std::numeric_limits<std::uint32_t>::max())
        throw std::length_error("Particle count exceeds kernel index range");
    check_params(params);

    // Every scene's grid exactly as PBFSolver::initialize builds it (the
    // bounds checks of spatial_grid.cu:11-27 included), laid end to end.
    const float h = params.smoothingRadius;
    std::vector<pbf_container> boxes(scenes.size());
    std::vector<std::size_t> first(scenes.size() + 1, 0);
    for (std::size_t k = 0; k < scenes.size(); ++k) {
        const BatchScene& sc = scenes[k];
        const float3 lo = sc.minBounds, hi = sc.maxBounds;
        check_container(lo, hi, params);
        if (!std::isfinite(lo.x) || !std::isfinite(lo.y) || !std::isfinite(lo.z) ||
            !std::isfinite(hi.x) || !std::isfinite(hi.y) || !std::isfinite(hi.z) ||
            hi.x <= lo.x || hi.y <= lo.y || hi.z <= lo.z)
            throw std::invalid_argument("Invalid spatial grid bounds");
        boxes[k] = pbf_container{lo, hi};
        first[k + 1] = first[k] + sc.particleCount;
    }
    std::size_t cells = 0;
    const std::vector<pbfint::SceneGeom> geom = scene_grids(boxes, first, h, &cells);
    std::vector<std::uint32_t> sceneOf(total);
    for (std::size_t k = 0; k < scenes.size(); ++k)
        for (std::size_t i = first[k]; i < first[k + 1]; ++i)
            sceneOf[i] = static_cast<std::uint32_t>(k);

    PBFSolver::State& s = *solver_.state_;
    GridCore grid = GridCore::createKeyed(total, cells, h);
    s.drop_graphs();
    s.maxParticles = 0;
    s.count = 0;
    s.sceneCount = 0;
    s.grid = std::move(grid);
    s.collision.setContainer(Container{scenes[0].minBounds, scenes[0].maxBounds});
    s.ensure_stream();
    s.allocate(total, stride);
    s.sceneGeom.allocate(scenes.size());
    s.sceneGeom.copyFromHostToDevice(geom.data(), geom.size());
    s.sceneContainers.allocate(scenes.size());
    s.sceneContainers.copyFromHostToDevice(boxes.data(), boxes.size());
    s.sceneOf.allocate(total);
    s.sceneOf.copyFromHostToDevice(sceneOf.data(), total);
    s.sceneFirst = std::move(first);
    s.sceneBoxes = std::move(boxes);
    s.sceneCount = scenes.size();
    s.params = params;
    s.make_skin_grid(total);
    s.maxParticles = total;
}

void PBFBatchSolver::setParticles(const float4* positions, const float4* velocities,
                                  std::size_t particleCount) {
    const PBFSolver::State& s = *solver_.state_;
    if (s.sceneCount == 0)
        throw std::logic_error("PBFBatchSolver must be initialized before setting particles");
    if (particleCount != s.maxParticles)
        throw std::out_of_range("A batch takes every scene's particles at once");
    solver_.setParticles(positions, velocities, particleCount);
}

void PBFBatchSolver::step() {
    if (solver_.state_->sceneCount == 0)
        throw std::logic_error("PBFBatchSolver must be initialized before stepping");
    solver_.step();
}

void PBFBatchSolver::run(std::size_t frameCount) {
    if (solver_.state_->sceneCount == 0)
        throw std::logic_error("PBFBatchSolver must be initialized before running");
    solver_.run(frameCount);
}

void PBFBatchSolver::copyPositionsToHost(float4* out, std::size_t particleCount) const {
    solver_.copyPositionsToHost(out, particleCount);
}

void PBFBatchSolver::copyVelocitiesToHost(float4* out, std::size_t particleCount) const {
    solver_.copyVelocitiesToHost(out, particleCount);
}

std::size_t PBFBatchSolver::sceneCount() const { return solver_.state_->sceneCount; }
std::size_t PBFBatchSolver::particleCount() const { return solver_.particleCount(); }

std::size_t PBFBatchSolver::sceneFirst(std::size_t scene) const {
    const PBFSolver::State& s = *solver_.state_;
    if (scene > s.sceneCount) throw std::out_of_range("Scene index exceeds the batch");
    return s.sceneFirst[scene];
}

void PBFBatchSolver::setGraphs(bool enabled) { solver_.setGraphs(enabled); }
void PBFBatchSolver::setIncrementalSort(bool enabled) { solver_.setIncrementalSort(enabled); }
void PBFBatchSolver::setFusedXsph(bool enabled) { solver_.setFusedXsph(enabled); }
void PBFBatchSolver::setNeighborSkin(float skin) { solver_.setNeighborSkin(skin); }
std::size_t PBFBatchSolver::staleFrames() const { return solver_.staleFrames(); }
void PBFBatchSolver::setTuning(const PBFSolver::Tuning& tuning) { solver_.setTuning(tuning); }
PBFSolver::Tuning PBFBatchSolver::tuning() const { return solver_.tuning(); }
std::vector<PBFSolver::TuningTime> PBFBatchSolver::autotune(
    std::size_t frames, const std::vector<PBFSolver::Tuning>& candidates) {
    if (solver_.state_->sceneCount == 0)
        throw std::logic_error("PBFBatchSolver must be initialized before autotuning");
    return solver_.autotune(frames, candidates);
}
void PBFBatchSolver::setProfiling(bool enabled) { solver_.setProfiling(enabled); }
std::vector<PBFSolver::KernelTime> PBFBatchSolver::kernelTimes() const {
    return solver_.kernelTimes();
}

// ---------------------------------------------------------------------------
// PBFAdjoint: checkpointed forward, taped replay, adjoint kernels
// ---------------------------------------------------------------------------

namespace {

// The working set at a frame boundary (slot order, with the slot ids).
struct Checkpoint {
    DeviceBuffer<float4> x, v;
    DeviceBuffer<std::uint32_t> id;
};

}  // namespace

struct PBFAdjoint::Impl {
    PBFSolver* solver = nullptr;
    std::size_t frames = 0;
    std::size_t every = 0;
    std::size_t n = 0;
    std::vector<Checkpoint> checkpoints;  // frame 0, every, 2 every, ...
    Checkpoint last;                      // the state forward() ended in
    std::vector<SubstepTape> tapes;
    // The segment whose substeps the tapes hold (forward() tapes its last
    // segment as it runs it; backward() ends holding segment 0); npos: none.
    std::size_t tapedSegment = static_cast<std::size_t>(-1);
    DeviceBuffer<float4> pb, vb, x0b, v1b, v3b, etab, wb, prb, gsum;
    DeviceBuffer<float> cterm, lamb;
    DeviceBuffer<double> acc;

    PBFSolver::State& S() { return *solver->state_; }

    void save(Checkpoint& c) {
        PBFSolver::State& s = S();
        c.x.allocate(n);
        c.v.allocate(n);
        c.id.allocate(n);
        s.tape_copy(c.x, s.x, n);
        s.tape_copy(c.v, s.v, n);
        s.tape_copy(c.id, s.id, n);
    }
    void restore(const Checkpoint& c) {
        PBFSolver::State& s = S();
        const auto copy = [&](auto& dst, const auto& src, std::size_t bytes) {
            hip_check(hipMemcpyAsync(dst.data(), src.data(), bytes, hipMemcpyDeviceToDevice,
                                     s.stream));
        };
        copy(s.x, c.x, n * sizeof(float4));
        copy(s.v, c.v, n * sizeof(float4));
        copy(s.id, c.id, n * sizeof(std::uint32_t));
        s.forget_order();
    }
};

PBFAdjoint::PBFAdjoint(PBFSolver& solver) : impl_(new Impl()) { impl_->solver = &solver; }
PBFAdjoint::~PBFAdjoint() = default;

namespace {
// Runs f when the scope ends.
template <class F>
class OnExit {
public:
    explicit OnExit(F f) : f_(f) {}
    ~OnExit() { f_(); }
    OnExit(const OnExit&) = delete;
    OnExit& operator=(const OnExit&) = delete;

private:
    F f_;
};
}  // namespace

void PBFAdjoint::forward(std::size_t frameCount, std::size_t checkpointEvery) {
    Impl& m = *impl_;
    PBFSolver::State& s = m.S();
    if (s.maxParticles == 0) throw std::logic_error("PBFSolver must be initialized before running");
    if (s.sceneCount != 0) throw std::invalid_argument("The adjoint takes a single-scene solver");
    if (checkpointEvery == 0) throw std::invalid_argument("checkpointEvery must be positive");
    m.frames = frameCount;
    m.every = checkpointEvery;
    m.n = s.count;
    m.checkpoints.clear();
    // The adjoint's tapes and reverse kernels follow the reference's five
    // grid builds per substep: no skin lists while it runs frames.
    const bool wasExact = s.adjointExact;
    s.adjointExact = true;
    const OnExit restore([&s, wasExact] { s.adjointExact = wasExact; });
    m.tapedSegment = static_cast<std::size_t>(-1);
    s.reset_status();
    // The last segment is taped as it runs: backward() starts with it, and
    // its replay would record exactly these substeps again.
    const std::size_t lastSegment = frameCount == 0 ? 0 : (frameCount - 1) / checkpointEvery;
    const std::size_t lastFirst = lastSegment * checkpointEvery;
    try {
        for (std::size_t f = 0; f < frameCount; ++f) {
            if (f % checkpointEvery == 0) {
                m.checkpoints.emplace_back();
                m.save(m.checkpoints.back());
            }
            if (f == lastFirst) {
                const std::size_t subs =
                    (frameCount - lastFirst) * static_cast<std::size_t>(s.params.substeps);
                if (m.tapes.size() < subs) m.tapes.resize(subs);
                for (std::size_t q = 0; q < subs; ++q)
                    m.tapes[q].allocate(m.n, s.nbr.size(), s.params.solverIterations);
                s.tapes = &m.tapes;
                s.tapeNext = 0;
            }
            s.frame();
        }
    } catch (...) {
        s.tapes = nullptr;
        throw;
    }
    s.tapes = nullptr;
    m.save(m.last);
    s.finish();
    if (frameCount != 0) m.tapedSegment = lastSegment;
}

void PBFAdjoint::backward(const float4* xBar, const float4* vBar, std::size_t particleCount,
                          float4* x0Bar, float4* v0Bar, double* viscosityBar,
                          double* vorticityBar) {
    Impl& m = *impl_;
    PBFSolver::State& s = m.S();
    if (m.checkpoints.empty() && m.frames != 0)
        throw std::logic_error("PBFAdjoint::backward needs a forward() first");
    if (particleCount != m.n || s.count != m.n)
        throw std::out_of_range("Requested adjoint count does not match active particles");
    const std::size_t n = m.n;
    const SimulationParams& P = s.params;
    const bool wasExact = s.adjointExact;
    s.adjointExact = true;
    const OnExit restore([&s, wasExact] { s.adjointExact = wasExact; });
    for (DeviceBuffer<float4>* b :
         {&m.pb, &m.vb, &m.x0b, &m.v1b, &m.v3b, &m.etab, &m.wb, &m.prb, &m.gsum})
        b->allocate(n);
    for (DeviceBuffer<float>* b : {&m.cterm, &m.lamb}) b->allocate(n);
    m.acc.allocate(2);
    hip_check(hipMemsetAsync(m.acc.data(), 0, 2 * sizeof(double), s.stream));
    const auto load = [&](DeviceBuffer<float4>& d, const float4* h) {
        if (h != nullptr)
            hip_check(hipMemcpyAsync(d.data(), h, n * sizeof(float4), hipMemcpyHostToDevice,
                                     s.stream));
        else
            hip_check(hipMemsetAsync(d.data(), 0, n * sizeof(float4), s.stream));
    };
    load(m.pb, xBar);
    load(m.vb, vBar);

    pbfint::AdjParams ap{};
    ap.h = P.smoothingRadius;
    ap.dt = P.dt / static_cast<float>(P.substeps);
    ap.visc = P.xsphViscosity;
    ap.vort = P.vorticityStrength;
    ap.radius = P.particleRadius;
    ap.restitution = P.collisionRestitution;
    ap.friction = P.collisionFriction;
    ap.mass = P.particleMass;
    ap.rho0 = P.restDensity;
    ap.lambdaEps = P.lambdaEpsilon;
    ap.scorrK = P.scorrK;
    ap.scorrN = P.scorrN;
    ap.scorrDq = P.scorrDeltaQ;
    ap.col = s.collision.state().view();
    const auto build = [&](SubstepTape& tp, int b) {
        return pbfint::AdjBuild{tp.nbr[b], tp.cnt[b], tp.ids[b], s.stride};
    };

    const std::size_t segments = m.checkpoints.size();
    for (std::size_t k = segments; k-- > 0;) {
        const std::size_t f0 = k * m.every;
        const std::size_t f1 = std::min(m.frames, f0 + m.every);
        const std::size_t subs = (f1 - f0) * static_cast<std::size_t>(P.substeps);
        if (k != m.tapedSegment) {
            // Replay the segment with the tape on.
            if (m.tapes.size() < subs) m.tapes.resize(subs);
            for (std::size_t q = 0; q < subs; ++q)
                m.tapes[q].allocate(n, s.nbr.size(), P.solverIterations);
            m.tapedSegment = static_cast<std::size_t>(-1);
            m.restore(m.checkpoints[k]);
            s.reset_status();
            s.tapes = &m.tapes;
            s.tapeNext = 0;
            try {
                for (std::size_t f = f0; f < f1; ++f) s.enqueue_frame();
            } catch (...) {
                s.tapes = nullptr;
                throw;
            }
            s.tapes = nullptr;
            s.finish();
            m.tapedSegment = k;
        }
        // Back through the segment's substeps.
        for (std::size_t q = subs; q-- > 0;) {
            SubstepTape& tp = m.tapes[q];
            pbfint::AdjBuffers w{m.pb.data(),   m.vb.data(),    m.x0b.data(),  m.v1b.data(),
                                 m.v3b.data(),  m.etab.data(),  m.wb.data(),   m.prb.data(),
                                 m.gsum.data(), m.cterm.data(), m.lamb.data(), m.acc.data()};
            const pbfint::AdjPost post{build(tp, P.solverIterations), tp.xs, tp.pf, tp.v3,
                                       tp.omega, tp.vin};
            launch_check(pbfint::adjoint_post(post, w, n, ap, s.stream));
            for (int it = P.solverIterations; it-- > 0;) {
                const pbfint::AdjIter iter{build(tp, it), tp.pk[it], tp.pre[it]};
                launch_check(pbfint::adjoint_iteration(iter, w, n, ap, s.stream));
            }
            launch_check(pbfint::adjoint_predict(w, n, ap.dt, s.stream));
            std::swap(m.pb, m.x0b);
            std::swap(m.vb, m.v1b);
        }
    }
    if (x0Bar != nullptr)
        hip_check(hipMemcpyAsync(x0Bar, m.pb.data(), n * sizeof(float4), hipMemcpyDeviceToHost,
                                 s.stream));
    if (v0Bar != nullptr)
        hip_check(hipMemcpyAsync(v0Bar, m.vb.data(), n * sizeof(float4), hipMemcpyDeviceToHost,
                                 s.stream));
    double acc[2] = {0.0, 0.0};
    hip_check(hipMemcpyAsync(acc, m.acc.data(), sizeof(acc), hipMemcpyDeviceToHost, s.stream));
    m.restore(m.last);
    hip_check(hipStreamSynchronize(s.stream));
    if (viscosityBar != nullptr) *viscosityBar = acc[0];
    if (vorticityBar != nullptr) *vorticityBar = acc[1];
}

}  // namespace pbf
