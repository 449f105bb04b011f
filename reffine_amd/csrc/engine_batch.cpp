// Batched and served execution.
//
// A batch runner collects work units -- one compiled query over its
// arguments (one record batch) each -- into a round and runs the round with
// one persistent launch per distinct query kernel over the list of its
// units' wave-tiles (rf_qb, see kernels/rf_device.h), one completion launch
// per kernel and one placement launch for selections. Everything a launch
// reads lives in buffers that are fixed per runner (a control region
// rewritten each round, a result region, a bump arena for the units'
// partials and selection buffers), so a round can be captured into a
// hipGraph once per set of kernels and replayed. Admission is by
// device-memory footprint: a unit that does not fit the arena (or a full
// unit table) is refused with RF_ERR_BUSY until the round is recycled.
//
// A group join unit runs its probe scan as an ordinary unit of the round
// (against the resident build side of its build table, shared with plain
// runs); once the round's scans are done, one post-phase per group join
// query concatenates all its units' joined rows (ranks offset by each
// unit's first group), sorts them once with a stable radix sort, finds the
// segment bounds and folds every unit's groups in one launch. A stable sort
// keeps each unit's rows in row order within a group, so every group folds
// in the same order as the unit's plain run.
//
// The serving loop keeps two runners on two streams: requests that arrive
// while one round is on the GPU fill the other, which launches as soon as it
// has work, so rounds overlap on the device (continuous batching).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

#include "engine.hpp"

namespace rf {

static constexpr int kMaxGroups = 8;     // distinct query kernels per round
static constexpr size_t kSlot = 512;     // result bytes per unit (Acc, count at the end)

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* need(size_t bytes) {
        if (bytes > cap || !p) {
            dev_free(p);
            p = nullptr;
            cap = 0;
            p = dev_alloc(std::max<size_t>(bytes, 256));
            cap = bytes;
        }
        return p;
    }
    ~DevBuf() { dev_free(p); }
};

}  // namespace rf

struct rf_batch {
    int cap = 0;
    bool use_graph = false;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;       // the round's scans (and results copied back)
    hipEvent_t post_done = nullptr;  // the group join post-phase
    // Device regions and their pinned host mirrors.
    char* arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
    char* d_ctl = nullptr;   // hdr[kMaxGroups] | start[cap + kMaxGroups] | work[cap]
    char* h_ctl = nullptr;
    size_t ctl_bytes = 0, off_start = 0, off_work = 0;
    char* d_res = nullptr;   // [cap][kSlot]
    char* h_res = nullptr;
    struct Unit {
        rf_query* q = nullptr;
        std::vector<rf_arg> args;
        rf::Launch l;
        std::vector<void*> outs;
        size_t footprint = 0;
        // Group join: the resident build side, the fold module and the
        // output columns (one row per qualifying build key).
        std::shared_ptr<rf::GjBuild> build;
        const rf::Spec* fold = nullptr;
        std::vector<void*> gj_outs;
        bool no_scan = false;  // no scan in the round (an empty group join, a plain unit)
        // A function the round's kernels do not take (run folds, computed
        // vectors): run on the plain path when added; its output kept here.
        bool plain = false;
        std::vector<char> plain_out;
        std::shared_ptr<rf_table> plain_tbl;
    };
    std::vector<Unit> units;
    int64_t join_groups = 0;  // groups of the round's group join units
    bool launched = false;
    bool post_pending = false;   // scans launched, post-phase not yet
    bool post_launched = false;
    std::map<std::vector<rf::Module*>, hipGraphExec_t> graphs;
    // Post-phase buffers (grow-only): concatenated ranks and payloads
    // (ping-pong), sort histogram, segment bounds, run descriptors.
    rf::DevBuf rank[2], pay[2][8], hist, seg, d_post;
    char* h_post = nullptr;
    size_t h_post_cap = 0;
    ~rf_batch();
};

rf_batch::~rf_batch() {
    if (launched && done) (void)hipEventSynchronize(done);
    if (post_launched && post_done) (void)hipEventSynchronize(post_done);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    rf::dev_free(arena);
    rf::dev_free(d_ctl);
    rf::dev_free(d_res);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_res) (void)hipHostFree(h_res);
    if (h_post) (void)hipHostFree(h_post);
    if (done) (void)hipEventDestroy(done);
    if (post_done) (void)hipEventDestroy(post_done);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace rf {

rf_batch* batch_create(int cap, int64_t arena_bytes, bool use_graph) {
    if (cap < 1 || cap > 65536) throw IRError("units per round must be 1..65536");
    if (arena_bytes < 4096) throw IRError("arena must hold at least 4 KiB");
    device();
    auto b = std::make_unique<rf_batch>();
    b->cap = cap;
    b->use_graph = use_graph;
    RF_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    RF_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    RF_HIP(hipEventCreateWithFlags(&b->post_done, hipEventDisableTiming));
    b->arena_bytes = (size_t)arena_bytes;
    b->arena = (char*)dev_alloc(b->arena_bytes);
    b->off_start = align256(sizeof(RfBatchHdr) * kMaxGroups);
    b->off_work = b->off_start + align256(8 * (size_t)(cap + kMaxGroups));
    b->ctl_bytes = b->off_work + sizeof(RfScanArgs) * (size_t)cap;
    b->d_ctl = (char*)dev_alloc(b->ctl_bytes);
    RF_HIP(hipHostMalloc((void**)&b->h_ctl, b->ctl_bytes, hipHostMallocDefault));
    std::memset(b->h_ctl, 0, b->ctl_bytes);
    b->d_res = (char*)dev_alloc(kSlot * (size_t)cap);
    RF_HIP(hipHostMalloc((void**)&b->h_res, kSlot * (size_t)cap, hipHostMallocDefault));
    return b.release();
}

static void* arena_take(rf_batch* b, size_t bytes, size_t& used) {
    const size_t n = align256(std::max<size_t>(bytes, 1));
    if (used + n > b->arena_bytes) return nullptr;
    void* p = b->arena + used;
    used += n;
    return p;
}

// What a unit launches: the query's own scan, or for a group join the probe
// scan against the resident build side (built now if not cached).
static void plan_unit(rf_batch::Unit& u, rf_query* q, const rf_arg* args, int64_t n_args) {
    u.q = q;
    u.args.assign(args, args + n_args);
    if (q->plan.seg || q->plan.runs || !q->pre.empty() || q->plan.sequential ||
        q->plan.source == Plan::MERGE || q->plan.composite) {
        // Run folds, computed vectors, sequential folds and union merges are
        // not batch kernels: the unit runs on the plain path when added
        // (batch_add).
        std::function<bool(const Type&)> holds_table = [&](const Type& t) {
            if (t.is_vector()) return true;
            for (auto& d : t.dtypes)
                if (holds_table(d)) return true;
            return false;
        };
        const Type& ot = q->plan.output_type;
        if (!ot.is_vector() && holds_table(ot))
            throw Unsupported("a struct output holding a table from a run fold in a batch");
        u.plain = true;
        u.no_scan = true;
        return;
    }
    // The query's spec cache (and a group join's build cache) may be used by
    // a plain run on another thread.
    std::lock_guard<std::mutex> lk(q->mu);
    if (!q->gj) {
        u.l = prepare_args(q, u.args.data(), n_args);
        return;
    }
    check_args(q->plan, u.args.data(), n_args);
    u.build = gj_build(q, u.args.data(), n_args);
    u.fold = &gj_fold_spec(q, u.args.data(), n_args);
    if (u.build->groups == 0) {
        u.no_scan = true;
        return;
    }
    rf_query* pq = gj_probe_query(q, *u.build);
    u.l = prepare_args(pq, u.args.data(), n_args);
    const RfProbeDev h = u.build->probe();
    const auto& probes = q->plan.gj->probe_plan.probes;
    for (size_t j = 0; j < probes.size(); j++)
        if (probes[j].external) u.l.args.probe[j] = h;
}

static size_t unit_footprint(const rf_batch::Unit& u) {
    if (u.plain) return 0;
    size_t f = 0;
    if (u.build) {
        if (u.no_scan) return 0;
        for (auto& t : u.q->plan.gj->out_types) f += align256((size_t)u.build->groups * t.width());
    }
    const KernelSource& ks = u.l.spec->ks;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;
    const int64_t nt = std::max<int64_t>(u.l.args.ntiles, 1);
    if (ks.acc_bytes) f += align256((size_t)nt * ks.acc_bytes);
    if (ks.has_sel) {
        const int64_t rows = std::max<int64_t>(u.l.end - u.l.lo, 1);
        f += 2 * align256((size_t)nt * 8);
        for (auto& t : ks.stage_types) f += align256((size_t)(nt * tile) * t.width());
        for (auto& t : ks.out_types) f += align256((size_t)rows * t.width());
    }
    return f;
}

// Device bytes a unit of `q` over `args` takes from a runner's arena.
int64_t batch_footprint(rf_query* q, const rf_arg* args, int64_t n_args) {
    rf_batch::Unit u;
    plan_unit(u, q, args, n_args);
    return (int64_t)unit_footprint(u);
}

// Adds a unit; returns false (nothing changed) when the round is full.
bool batch_add(rf_batch* b, rf_query* q, const rf_arg* args, int64_t n_args, int32_t* unit) {
    if (b->launched) throw IRError("the round is launched: wait for it and reset the batch");
    if ((int)b->units.size() >= b->cap) return false;
    rf_batch::Unit u;
    plan_unit(u, q, args, n_args);
    const int idx = (int)b->units.size();
    if (u.plain) {
        // Room for the output struct (or the table handle an Op output writes).
        size_t sz = sizeof(void*), al = 0;
        if (!q->plan.output_type.is_vector()) c_layout(q->plan.output_type, sz, al);
        u.plain_out.assign(std::max(sz, sizeof(void*)), 0);
        rf_table* t = nullptr;
        run_query(q, args, n_args, u.plain_out.data(), &t);
        u.plain_tbl.reset(t, [](rf_table* x) { table_free(x); });
    }
    if (u.no_scan) {
        b->units.push_back(std::move(u));
        *unit = idx;
        return true;
    }
    if (u.build && b->join_groups + u.build->groups >= ((int64_t)1 << 31)) return false;
    const KernelSource& ks = u.l.spec->ks;
    if ((size_t)ks.acc_bytes > kSlot - 8)
        throw Unsupported("reduction state too large for a batch result slot");
    RfScanArgs& A = u.l.args;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;
    const int64_t nt = std::max<int64_t>(A.ntiles, 1);
    size_t used = b->arena_used;
    char* res = b->d_res + kSlot * (size_t)idx;
    A.result = res;
    A.out_count = (int64_t*)(res + kSlot - 8);
    A.counters = nullptr;
    if (ks.acc_bytes && !(A.partials = arena_take(b, (size_t)nt * ks.acc_bytes, used)))
        return false;
    if (ks.has_sel) {
        const int64_t rows = std::max<int64_t>(u.l.end - u.l.lo, 1);
        A.wave_cnt = (int64_t*)arena_take(b, (size_t)nt * 8, used);
        A.wave_dst = (int64_t*)arena_take(b, (size_t)nt * 8, used);
        if (!A.wave_cnt || !A.wave_dst) return false;
        A.stage_rows = tile;
        A.nwaves = A.ntiles;
        A.out_cap = rows;
        for (size_t j = 0; j < ks.stage_types.size(); j++) {
            A.stage[j] = arena_take(b, (size_t)(nt * tile) * ks.stage_types[j].width(), used);
            if (!A.stage[j]) return false;
        }
        for (size_t j = 0; j < ks.out_types.size(); j++) {
            A.out[j] = arena_take(b, (size_t)rows * ks.out_types[j].width(), used);
            if (!A.out[j]) return false;
            u.outs.push_back(A.out[j]);
        }
    }
    if (u.build) {
        for (auto& t : q->plan.gj->out_types) {
            void* p = arena_take(b, (size_t)u.build->groups * t.width(), used);
            if (!p) return false;
            u.gj_outs.push_back(p);
        }
        b->join_groups += u.build->groups;
    }
    u.footprint = used - b->arena_used;
    b->arena_used = used;
    b->units.push_back(std::move(u));
    *unit = idx;
    return true;
}

// Launch group of a round: units sharing a kernel module, in a canonical
// (pointer) order so the same kernel set replays the same captured graph.
struct Group {
    Module* mod;
    int block;
    bool has_sel;
    std::vector<int> units;
};

static std::vector<Group> groups_of(rf_batch* b) {
    std::map<Module*, Group> m;
    for (size_t i = 0; i < b->units.size(); i++) {
        const auto& u = b->units[i];
        if (u.no_scan) continue;
        Module* mod = u.l.spec->mod;
        auto it = m.find(mod);
        if (it == m.end())
            it = m.emplace(mod, Group{mod, u.l.spec->ks.tune.block, u.l.spec->ks.has_sel, {}}).first;
        it->second.units.push_back((int)i);
    }
    if (m.size() > (size_t)kMaxGroups) throw Unsupported("more than 8 distinct query kernels in a round");
    std::vector<Group> g;
    for (auto& kv : m) g.push_back(kv.second);
    return g;
}

// The launches of a round on `s` (eager, or while capturing a graph).
static void enqueue_round(rf_batch* b, const std::vector<Group>& gs, bool full_copy) {
    hipStream_t s = b->stream;
    Device& d = device();
    // Control region: headers, item starts and the units' arguments.
    size_t bytes = b->ctl_bytes;
    if (!full_copy) bytes = b->off_work + sizeof(RfScanArgs) * b->units.size();
    RF_HIP(hipMemcpyAsync(b->d_ctl, b->h_ctl, bytes, hipMemcpyHostToDevice, s));
    for (size_t g = 0; g < gs.size(); g++) {
        RfBatchArgs B;
        B.hdr = (const RfBatchHdr*)b->d_ctl + g;
        B.work = (const RfScanArgs*)(b->d_ctl + b->off_work);
        B.item_start = (const int64_t*)(b->d_ctl + b->off_start);
        void* params[] = {&B};
        const Group& G = gs[g];
        const int grid = d.num_cus * G.mod->bscan_blocks_per_cu;
        RF_HIP(hipModuleLaunchKernel(G.mod->bscan, grid, 1, 1, G.block, 1, 1, 0, s, params, nullptr));
    }
    for (size_t g = 0; g < gs.size(); g++) {
        RfBatchArgs B;
        B.hdr = (const RfBatchHdr*)b->d_ctl + g;
        B.work = (const RfScanArgs*)(b->d_ctl + b->off_work);
        B.item_start = (const int64_t*)(b->d_ctl + b->off_start);
        void* params[] = {&B};
        const Group& G = gs[g];
        const int units = full_copy ? b->cap : (int)G.units.size();
        RF_HIP(hipModuleLaunchKernel(G.mod->bdone, units, 1, 1, G.block, 1, 1, 0, s, params, nullptr));
        if (G.has_sel)
            RF_HIP(hipModuleLaunchKernel(G.mod->bgather, d.num_cus * 4, 1, 1, 256, 1, 1, 0, s,
                                         params, nullptr));
    }
    const size_t rb = full_copy ? kSlot * (size_t)b->cap : kSlot * b->units.size();
    RF_HIP(hipMemcpyAsync(b->h_res, b->d_res, rb, hipMemcpyDeviceToHost, s));
}

void batch_launch(rf_batch* b) {
    if (b->launched) throw IRError("the round is already launched");
    b->post_pending = b->post_launched = false;
    for (auto& u : b->units)
        if (u.build && !u.no_scan) b->post_pending = true;
    std::vector<Group> gs = groups_of(b);
    if (gs.empty()) {
        b->launched = true;
        RF_HIP(hipEventRecord(b->done, b->stream));
        return;
    }
    // Host mirror of the control region: per group its header and item
    // starts; the units' arguments in group order.
    RfBatchHdr* hdr = (RfBatchHdr*)b->h_ctl;
    int64_t* start = (int64_t*)(b->h_ctl + b->off_start);
    RfScanArgs* work = (RfScanArgs*)(b->h_ctl + b->off_work);
    std::memset(hdr, 0, sizeof(RfBatchHdr) * kMaxGroups);
    int64_t ub = 0, sb = 0;
    for (size_t g = 0; g < gs.size(); g++) {
        hdr[g].nunits = (int64_t)gs[g].units.size();
        hdr[g].unit_base = ub;
        hdr[g].start_base = sb;
        int64_t items = 0;
        for (int ui : gs[g].units) {
            start[sb++] = items;
            items += b->units[ui].l.args.ntiles;
            work[ub++] = b->units[ui].l.args;
        }
        start[sb++] = items;
        hdr[g].nitems = items;
    }
    if (b->use_graph) {
        std::vector<Module*> key;
        for (auto& g : gs) key.push_back(g.mod);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t graph;
            RF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_round(b, gs, true);
            } catch (...) {
                (void)hipStreamEndCapture(b->stream, &graph);
                throw;
            }
            RF_HIP(hipStreamEndCapture(b->stream, &graph));
            hipGraphExec_t exec;
            RF_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            it = b->graphs.emplace(key, exec).first;
        }
        RF_HIP(hipGraphLaunch(it->second, b->stream));
    } else {
        enqueue_round(b, gs, false);
    }
    RF_HIP(hipEventRecord(b->done, b->stream));
    b->launched = true;
}

// The group join post-phase of a finished round (joined row counts are in
// the host copy of the result slots). Per fold module (one group join query
// and nullability): concatenate the units' joined rows, one stable radix
// sort by (unit's first group + rank), segment bounds, one fold launch over
// every unit's groups.
static void launch_post(rf_batch* b) {
    hipStream_t s = b->stream;
    std::map<const Spec*, std::vector<int>> by_fold;
    for (size_t i = 0; i < b->units.size(); i++) {
        const auto& u = b->units[i];
        if (u.build && !u.no_scan) by_fold[u.fold].push_back((int)i);
    }
    // Descriptors per fold module: segs[n] | folds[n] | gstart[n + 1].
    size_t bytes = 0;
    for (auto& kv : by_fold) {
        const size_t n = kv.second.size();
        bytes += align256(n * sizeof(RfCatSeg)) + align256(n * sizeof(RfFoldArgs)) +
                 align256((n + 1) * 8);
    }
    if (bytes > b->h_post_cap) {
        if (b->h_post) RF_HIP(hipHostFree(b->h_post));
        b->h_post = nullptr;
        b->h_post_cap = 0;
        RF_HIP(hipHostMalloc((void**)&b->h_post, bytes, hipHostMallocDefault));
        b->h_post_cap = bytes;
    }
    char* dpost = (char*)b->d_post.need(bytes);
    // Joined rows and groups per fold module; the shared sort and segment
    // buffers are sized for the largest before anything is enqueued (a
    // buffer reallocated between two modules' launches would be freed while
    // the first module's launches still use it).
    std::vector<int64_t> rows_of, groups_of_fold;
    int64_t max_rows = 0, max_groups = 0;
    int max_np = 0;
    for (auto& kv : by_fold) {
        int64_t jt = 0, gt = 0;
        for (int ui : kv.second) {
            const rf_batch::Unit& u = b->units[ui];
            jt += u.l.empty ? 0 : *(const int64_t*)(b->h_res + kSlot * (size_t)ui + kSlot - 8);
            gt += u.build->groups;
        }
        rows_of.push_back(jt);
        groups_of_fold.push_back(gt);
        max_rows = std::max(max_rows, jt);
        max_groups = std::max(max_groups, gt);
        max_np = std::max(max_np, (int)b->units[kv.second[0]].q->plan.gj->pay_types.size());
    }
    if (max_rows > 0) {
        for (int x = 0; x < 2; x++) {
            b->rank[x].need((size_t)max_rows * 4);
            for (int j = 0; j < max_np; j++) b->pay[x][j].need((size_t)max_rows * 8);
        }
        b->hist.need((size_t)aot_sort_hist_entries(max_rows) * 4);
    }
    b->seg.need((size_t)max_groups * 16);
    size_t off = 0, fi = 0;
    for (auto& kv : by_fold) {
        const std::vector<int>& us = kv.second;
        const rf_batch::Unit& u0 = b->units[us[0]];
        const GroupJoinPlan& gp = *u0.q->plan.gj;
        const int np = (int)gp.pay_types.size();
        const size_t n = us.size();
        const size_t o_seg = off, o_fold = o_seg + align256(n * sizeof(RfCatSeg)),
                     o_gs = o_fold + align256(n * sizeof(RfFoldArgs));
        off = o_gs + align256((n + 1) * 8);
        RfCatSeg* segs = (RfCatSeg*)(b->h_post + o_seg);
        RfFoldArgs* folds = (RfFoldArgs*)(b->h_post + o_fold);
        int64_t* gstart = (int64_t*)(b->h_post + o_gs);
        const int64_t jt = rows_of[fi], gt = groups_of_fold[fi];
        fi++;
        int64_t dst = 0, g0 = 0, maxj = 0;
        for (size_t k = 0; k < n; k++) {
            const rf_batch::Unit& u = b->units[us[k]];
            const int64_t J =
                u.l.empty ? 0 : *(const int64_t*)(b->h_res + kSlot * (size_t)us[k] + kSlot - 8);
            std::memset(&segs[k], 0, sizeof(RfCatSeg));
            segs[k].rank = (const uint32_t*)u.outs[0];
            for (int j = 0; j < np; j++) segs[k].pay[j] = u.outs[1 + j];
            segs[k].n = J;
            segs[k].dst = dst;
            segs[k].base = (uint32_t)g0;
            gstart[k] = g0;
            dst += J;
            g0 += u.build->groups;
            maxj = std::max(maxj, J);
        }
        gstart[n] = g0;
        void* kb[2] = {b->rank[0].p, b->rank[1].p};
        void* pb[2][8] = {};
        int32_t pw[8] = {};
        for (int j = 0; j < np; j++) {
            pw[j] = gp.pay_types[j].width();
            pb[0][j] = b->pay[0][j].p;
            pb[1][j] = b->pay[1][j].p;
        }
        int64_t* lo = (int64_t*)b->seg.p;
        int64_t* hi = lo + gt;
        int w = 0;
        if (jt > 0) {
            RF_HIP(hipMemcpyAsync(dpost + o_seg, b->h_post + o_seg, n * sizeof(RfCatSeg),
                                  hipMemcpyHostToDevice, s));
            RfCatArgs CA;
            std::memset(&CA, 0, sizeof CA);
            CA.rank = (uint32_t*)kb[0];
            for (int j = 0; j < np; j++) {
                CA.pay[j] = pb[0][j];
                CA.width[j] = pw[j];
            }
            CA.np = np;
            RF_HIP(aot_gj_concat((const RfCatSeg*)(dpost + o_seg), (int32_t)n, maxj, CA, s));
            if (jt > 1)
                RF_HIP(aot_radix_sort(kb, pb, pw, np, 4, jt, bits_of((uint64_t)(gt - 1)), 0,
                                      (uint32_t*)b->hist.p, &w, s));
        }
        RF_HIP(aot_fill_u64(lo, 0, 2 * gt, s));
        if (jt > 0) RF_HIP(aot_seg_bounds((const uint32_t*)kb[w], jt, lo, hi, s));
        for (size_t k = 0; k < n; k++) {
            const rf_batch::Unit& u = b->units[us[k]];
            RfFoldArgs& A = folds[k];
            gj_fold_args(u.q, *u.fold, *u.build, u.args.data(), (int64_t)u.args.size(), A);
            A.seg_lo = lo + gstart[k];
            A.seg_hi = hi + gstart[k];
            for (int j = 0; j < np; j++) A.pay[j] = pb[w][j];
            for (size_t j = 0; j < u.gj_outs.size(); j++) A.out[j] = u.gj_outs[j];
        }
        RF_HIP(hipMemcpyAsync(dpost + o_fold, b->h_post + o_fold, o_gs + (n + 1) * 8 - o_fold,
                              hipMemcpyHostToDevice, s));
        RfFoldBatch FB;
        FB.list = (const RfFoldArgs*)(dpost + o_fold);
        FB.gstart = (const int64_t*)(dpost + o_gs);
        FB.n = (int64_t)n;
        void* params[] = {&FB};
        if (jt >= kWaveFoldRows * gt) {  // large groups: a wave per group
            const int g = (int)std::min<int64_t>((gt + 3) / 4, 65536);
            RF_HIP(hipModuleLaunchKernel(kv.first->mod->fold_wb, g, 1, 1, 256, 1, 1, 0, s, params,
                                         nullptr));
        } else {
            const int g = (int)std::min<int64_t>((gt + 255) / 256, 16384);
            RF_HIP(hipModuleLaunchKernel(kv.first->mod->fold_b, g, 1, 1, 256, 1, 1, 0, s, params,
                                         nullptr));
        }
    }
    RF_HIP(hipEventRecord(b->post_done, s));
    b->post_pending = false;
    b->post_launched = true;
}

bool batch_poll(rf_batch* b) {
    if (!b->launched) return false;
    if (b->post_pending) {
        hipError_t e = hipEventQuery(b->done);
        if (e == hipErrorNotReady) return false;
        RF_HIP(e);
        launch_post(b);
    }
    hipError_t e = hipEventQuery(b->post_launched ? b->post_done : b->done);
    if (e == hipErrorNotReady) return false;
    RF_HIP(e);
    return true;
}

void batch_wait(rf_batch* b) {
    if (!b->launched) throw IRError("no round launched");
    RF_HIP(hipEventSynchronize(b->done));
    if (b->post_pending) launch_post(b);
    if (b->post_launched) RF_HIP(hipEventSynchronize(b->post_done));
}

// Output of a unit after the round finished. A selection result (and a
// group join's table) is a view of the runner's arena: valid until the
// runner is reset.
void batch_result(rf_batch* b, int32_t unit, void* output_addr, rf_table** out_table) {
    if (!b->launched) throw IRError("no round launched");
    if (b->post_pending) throw IRError("the round has not finished");
    if (unit < 0 || unit >= (int32_t)b->units.size()) throw IRError("no such unit");
    auto& u = b->units[unit];
    if (u.plain) {
        if (u.q->plan.output_type.is_vector()) {
            // The unit's table stays owned by the batch (plain_tbl); the caller
            // gets a view sharing its buffers (valid after the reset), never
            // the internal handle.
            if (!out_table) throw IRError("an Op output needs out_table");
            rf_table* t = u.plain_tbl ? new rf_table(*u.plain_tbl) : nullptr;
            *out_table = t;
            if (output_addr) *(void**)output_addr = t;
            return;
        }
        if (output_addr) std::memcpy(output_addr, u.plain_out.data(), u.plain_out.size());
        if (out_table) *out_table = nullptr;
        return;
    }
    if (u.build) {
        if (!out_table) throw IRError("an Op output needs out_table");
        const GroupJoinPlan& gp = *u.q->plan.gj;
        rf_table* t = u.no_scan ? alloc_table(gp.out_types, 0, 1)
                                : table_view(gp.out_types, u.gj_outs, u.build->groups, 1);
        *out_table = t;
        if (output_addr) *(void**)output_addr = t;
        return;
    }
    const KernelSource& ks = u.l.spec->ks;
    const char* res = b->h_res + kSlot * (size_t)unit;
    rf_table* sel = nullptr;
    if (ks.has_sel) {
        const int64_t count = u.l.empty ? 0 : *(const int64_t*)(res + kSlot - 8);
        sel = table_view(ks.out_types, u.outs, count, u.q->plan.ops[0].n_iters);
    }
    finish_output(u.q, ks, u.args.data(), (int64_t)u.args.size(), res, u.l.empty, sel,
                  output_addr, out_table);
    if (sel && !out_table) table_free(sel);
}

void batch_reset(rf_batch* b) {
    if (b->launched) {
        RF_HIP(hipEventSynchronize(b->done));
        if (b->post_launched) RF_HIP(hipEventSynchronize(b->post_done));
    }
    b->launched = false;
    b->post_pending = b->post_launched = false;
    b->units.clear();
    b->arena_used = 0;
    b->join_groups = 0;
}

int batch_units(const rf_batch* b) { return (int)b->units.size(); }

// ------------------------------------------------------------ serving

static double now_us(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

// Copies the first column of a request's selection (or group join table)
// to its host buffer, up to its capacity.
static void copy_first_column(rf_request& rq, const rf_table* sel) {
    rq.out_rows = sel ? sel->rows() : 0;
    if (sel && rq.sel_out && rq.out_rows > 0 && !sel->cols.empty()) {
        const Column& c0 = sel->cols[0];
        const size_t nb =
            std::min<size_t>((size_t)rq.out_rows * c0.type.width(), (size_t)rq.sel_cap);
        RF_HIP(hipMemcpy(rq.sel_out, c0.values, nb, hipMemcpyDeviceToHost));
    }
}

// One request through the plain path (rf_query_run); fills its outputs and
// status (not its completion time).
static void run_plain_request(rf_request& rq) {
    rf_table* sel = nullptr;
    rq.status = 0;
    rq.out_rows = 0;
    try {
        run_query(rq.q, rq.args, rq.n_args, rq.output_addr, &sel);
        copy_first_column(rq, sel);
    } catch (const Unsupported&) {
        rq.status = RF_ERR_UNSUPPORTED;
    } catch (const std::exception&) {
        rq.status = RF_ERR_INVALID;
    }
    if (sel) table_free(sel);
}

// Runners kept between serving calls (their arenas, result regions,
// post-phase buffers and captured graphs), by (units, arena, graph): a
// serving loop started again does not allocate or capture again. At most
// kPoolRunners are kept; the rest are freed when returned.
static constexpr size_t kPoolRunners = 4;
static std::mutex g_pool_mu;
// Never destroyed: runners hold HIP objects, which must not be released
// from static destructors after the runtime has shut down.
static auto& g_pool = *new std::vector<std::unique_ptr<rf_batch>>();

static std::unique_ptr<rf_batch> runner_take(int32_t units, int64_t arena, bool graph) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (size_t i = 0; i < g_pool.size(); i++) {
            rf_batch* b = g_pool[i].get();
            if (b->cap == units && (int64_t)b->arena_bytes == arena && b->use_graph == graph) {
                std::unique_ptr<rf_batch> r = std::move(g_pool[i]);
                g_pool.erase(g_pool.begin() + (std::ptrdiff_t)i);
                return r;
            }
        }
    }
    return std::unique_ptr<rf_batch>(batch_create(units, arena, graph));
}

struct RunnerLease {
    std::unique_ptr<rf_batch> b;
    ~RunnerLease() {
        if (!b) return;
        try {
            batch_reset(b.get());
        } catch (...) {
            return;  // a runner in a bad state is freed, not pooled
        }
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (g_pool.size() < kPoolRunners) g_pool.push_back(std::move(b));
    }
};

// Runs `n` requests (arrival times in us from the start, ascending) through
// `runners` runners (2..4; 0: 3) with continuous batching: arrivals are
// admitted into an idle runner, which launches at once, so up to `runners`
// rounds are in flight on their streams. Fills each request's completion
// time, status and outputs. Returns the wall time in us.
double serve(rf_request* reqs, int64_t n, int32_t max_units, int64_t arena_bytes, bool use_graph,
             int32_t n_runners) {
    for (int64_t i = 1; i < n; i++)
        if (reqs[i].arrival_us < reqs[i - 1].arrival_us)
            throw IRError("requests must be ordered by arrival time");
    if (max_units < 1 || max_units > 65536) throw IRError("units per round must be 1..65536");
    if (arena_bytes < 4096) throw IRError("arena must hold at least 4 KiB");
    // Three rounds in flight served the 4096-query mix 20 % faster than two
    // (151-158K against 123-131K queries/s closed loop); four were slower.
    if (n_runners == 0) n_runners = 3;
    if (n_runners < 2 || n_runners > (int32_t)kPoolRunners) throw IRError("runners must be 2..4");
    const int R = n_runners;
    std::vector<RunnerLease> leases(R);
    std::vector<rf_batch*> runners(R);
    for (int k = 0; k < R; k++) {
        leases[k].b = runner_take(max_units, arena_bytes, use_graph);
        runners[k] = leases[k].b.get();
    }
    std::vector<std::vector<int64_t>> members(R);  // request index per unit
    int64_t next = 0, finished = 0;
    int filling = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (finished < n) {
        bool progress = false;
        // Collect finished rounds.
        for (int k = 0; k < R; k++) {
            rf_batch* b = runners[k];
            if (!b->launched || !batch_poll(b)) continue;
            const double t = now_us(t0);
            for (size_t i = 0; i < members[k].size(); i++) {
                rf_request& rq = reqs[members[k][i]];
                rf_table* sel = nullptr;
                rq.status = 0;
                try {
                    batch_result(b, (int32_t)i, rq.output_addr, &sel);
                    copy_first_column(rq, sel);
                } catch (...) {
                    rq.status = RF_ERR_INVALID;
                }
                if (sel) table_free(sel);
                rq.done_us = t;
                finished++;
            }
            members[k].clear();
            batch_reset(b);
            progress = true;
        }
        // Admit arrived requests into the filling runner: the next idle one
        // after the last launched.
        for (int k = 0; k < R && runners[filling]->launched; k++) filling = (filling + 1) % R;
        const double t = now_us(t0);
        rf_batch* f = runners[filling];
        if (!f->launched) {
            while (next < n && reqs[next].arrival_us <= t) {
                rf_request& rq = reqs[next];
                int32_t unit = -1;
                bool added = false;
                int status = 0;
                try {
                    added = batch_add(f, rq.q, rq.args, rq.n_args, &unit);
                    if (!added && batch_units(f) == 0) status = RF_ERR_NOMEM;  // never fits
                } catch (const Unsupported&) {
                    status = RF_ERR_UNSUPPORTED;
                } catch (const std::exception&) {
                    status = RF_ERR_INVALID;
                }
                if (status) {
                    rq.status = status;
                    rq.out_rows = 0;
                    rq.done_us = now_us(t0);
                    finished++;
                    next++;
                    progress = true;
                    continue;
                }
                if (!added) break;  // the round is full: wait for a runner
                members[filling].push_back(next);
                next++;
                progress = true;
            }
            if (batch_units(f) > 0) {
                batch_launch(f);
                filling = (filling + 1) % R;
                progress = true;
            }
        }
        if (!progress) std::this_thread::yield();
    }
    return now_us(t0);
}

// The plain counterpart of serve(): each request, once it has arrived, runs
// alone through rf_query_run (one query per launch sequence, synchronous).
double serve_plain(rf_request* reqs, int64_t n) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t i = 0; i < n; i++) {
        rf_request& rq = reqs[i];
        while (now_us(t0) < rq.arrival_us) std::this_thread::yield();
        run_plain_request(rq);
        rq.done_us = now_us(t0);
    }
    return now_us(t0);
}

}  // namespace rf
