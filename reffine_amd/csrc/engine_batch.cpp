// Batched and served execution.
//
// A batch runner collects work units -- one compiled query over its
// arguments (one record batch) each -- into a round and runs the round with
// one persistent launch per distinct query kernel over the list of its
// units' wave-tiles (rf_qb, see kernels/rf_device.h), one completion launch
// per kernel and one placement launch for selections. Everything a launch reads lives in buffers that are
// fixed per runner (a control region rewritten each round, a result region,
// a bump arena for the units' partials and selection buffers), so a round can
// be captured into a hipGraph once per set of kernels and replayed.
// Admission is by device-memory footprint: a unit that does not fit the
// arena (or a full unit table) is refused with RF_ERR_BUSY until the round
// is recycled.
//
// The serving loop keeps two runners on two streams: requests that arrive
// while one round is on the GPU fill the other, which launches as soon as it
// has work, so rounds overlap on the device (continuous batching).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "engine.hpp"

namespace rf {

static constexpr int kMaxGroups = 8;     // distinct query kernels per round
static constexpr size_t kSlot = 512;     // result bytes per unit (Acc, count at the end)

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace rf

struct rf_batch {
    int cap = 0;
    bool use_graph = false;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
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
    };
    std::vector<Unit> units;
    bool launched = false;
    std::map<std::vector<rf::Module*>, hipGraphExec_t> graphs;
    ~rf_batch();
};

rf_batch::~rf_batch() {
    if (launched && done) (void)hipEventSynchronize(done);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    rf::dev_free(arena);
    rf::dev_free(d_ctl);
    rf::dev_free(d_res);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_res) (void)hipHostFree(h_res);
    if (done) (void)hipEventDestroy(done);
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

// Device bytes a unit of `q` over `args` takes from a runner's arena.
int64_t batch_footprint(rf_query* q, const rf_arg* args, int64_t n_args) {
    if (q->gj) throw Unsupported("group joins do not run in a batch");
    Launch l = prepare_args(q, args, n_args);
    const KernelSource& ks = l.spec->ks;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;
    const int64_t nt = std::max<int64_t>(l.args.ntiles, 1);
    size_t f = 0;
    if (ks.acc_bytes) f += align256((size_t)nt * ks.acc_bytes);
    if (ks.has_sel) {
        const int64_t rows = std::max<int64_t>(l.end - l.lo, 1);
        f += 2 * align256((size_t)nt * 8);
        for (auto& t : ks.stage_types) f += align256((size_t)(nt * tile) * t.width());
        for (auto& t : ks.out_types) f += align256((size_t)rows * t.width());
    }
    return (int64_t)f;
}

// Adds a unit; returns false (nothing changed) when the round is full.
bool batch_add(rf_batch* b, rf_query* q, const rf_arg* args, int64_t n_args, int32_t* unit) {
    if (b->launched) throw IRError("the round is launched: wait for it and reset the batch");
    if (q->gj) throw Unsupported("group joins do not run in a batch");
    if ((int)b->units.size() >= b->cap) return false;
    rf_batch::Unit u;
    u.q = q;
    u.args.assign(args, args + n_args);
    u.l = prepare_args(q, u.args.data(), n_args);
    const KernelSource& ks = u.l.spec->ks;
    if ((size_t)ks.acc_bytes > kSlot - 8)
        throw Unsupported("reduction state too large for a batch result slot");
    RfScanArgs& A = u.l.args;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;
    const int64_t nt = std::max<int64_t>(A.ntiles, 1);
    size_t used = b->arena_used;
    const int idx = (int)b->units.size();
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
    if (b->units.empty()) {
        b->launched = true;
        RF_HIP(hipEventRecord(b->done, b->stream));
        return;
    }
    std::vector<Group> gs = groups_of(b);
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

bool batch_poll(rf_batch* b) {
    if (!b->launched) return false;
    hipError_t e = hipEventQuery(b->done);
    if (e == hipErrorNotReady) return false;
    RF_HIP(e);
    return true;
}

void batch_wait(rf_batch* b) {
    if (!b->launched) throw IRError("no round launched");
    RF_HIP(hipEventSynchronize(b->done));
}

// Output of a unit after the round finished. A selection result is a view
// of the runner's arena: valid until the runner is reset.
void batch_result(rf_batch* b, int32_t unit, void* output_addr, rf_table** out_table) {
    if (!b->launched) throw IRError("no round launched");
    if (unit < 0 || unit >= (int32_t)b->units.size()) throw IRError("no such unit");
    auto& u = b->units[unit];
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
    if (b->launched) RF_HIP(hipEventSynchronize(b->done));
    b->launched = false;
    b->units.clear();
    b->arena_used = 0;
}

int batch_units(const rf_batch* b) { return (int)b->units.size(); }

// ------------------------------------------------------------ serving

static double now_us(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

// One request through the plain path (rf_query_run); fills its outputs and
// status (not its completion time).
static void run_plain_request(rf_request& rq) {
    rf_table* sel = nullptr;
    rq.status = 0;
    rq.out_rows = 0;
    try {
        run_query(rq.q, rq.args, rq.n_args, rq.output_addr, &sel);
        rq.out_rows = sel ? sel->rows() : 0;
        if (sel && rq.sel_out && rq.out_rows > 0 && !sel->cols.empty()) {
            const Column& c0 = sel->cols[0];
            const size_t nb = std::min<size_t>((size_t)rq.out_rows * c0.type.width(),
                                               (size_t)rq.sel_cap);
            RF_HIP(hipMemcpy(rq.sel_out, c0.values, nb, hipMemcpyDeviceToHost));
        }
    } catch (const Unsupported&) {
        rq.status = RF_ERR_UNSUPPORTED;
    } catch (const std::exception&) {
        rq.status = RF_ERR_INVALID;
    }
    if (sel) table_free(sel);
}

// Runs `n` requests (arrival times in us from the start, ascending) through
// two runners with continuous batching; fills each request's completion
// time, status and outputs. Requests a runner cannot take (group joins) run
// on a worker thread through the plain path on the library stream, so they
// overlap with the rounds on the runners' streams. Returns the wall time in
// us.
double serve(rf_request* reqs, int64_t n, int32_t max_units, int64_t arena_bytes, bool use_graph) {
    for (int64_t i = 1; i < n; i++)
        if (reqs[i].arrival_us < reqs[i - 1].arrival_us)
            throw IRError("requests must be ordered by arrival time");
    std::unique_ptr<rf_batch> r0(batch_create(max_units, arena_bytes, use_graph));
    std::unique_ptr<rf_batch> r1(batch_create(max_units, arena_bytes, use_graph));
    rf_batch* runners[2] = {r0.get(), r1.get()};
    std::vector<std::vector<int64_t>> members(2);  // request index per unit
    int64_t next = 0, finished = 0;
    int filling = 0;
    const auto t0 = std::chrono::steady_clock::now();
    // The worker for requests that run alone.
    std::mutex wmu;
    std::condition_variable wcv;
    std::deque<int64_t> wq;
    bool wstop = false;
    std::atomic<int64_t> wdone{0};
    std::thread worker([&] {
        for (;;) {
            int64_t i;
            {
                std::unique_lock<std::mutex> lk(wmu);
                wcv.wait(lk, [&] { return wstop || !wq.empty(); });
                if (wq.empty()) return;
                i = wq.front();
                wq.pop_front();
            }
            run_plain_request(reqs[i]);
            reqs[i].done_us = now_us(t0);
            wdone.fetch_add(1);
        }
    });
    struct Joiner {
        std::thread& t;
        std::mutex& m;
        std::condition_variable& cv;
        bool& stop;
        ~Joiner() {
            {
                std::lock_guard<std::mutex> lk(m);
                stop = true;
            }
            cv.notify_all();
            if (t.joinable()) t.join();
        }
    } joiner{worker, wmu, wcv, wstop};
    while (finished + wdone.load() < n) {
        bool progress = false;
        // Collect finished rounds.
        for (int k = 0; k < 2; k++) {
            rf_batch* b = runners[k];
            if (!b->launched || !batch_poll(b)) continue;
            const double t = now_us(t0);
            for (size_t i = 0; i < members[k].size(); i++) {
                rf_request& rq = reqs[members[k][i]];
                rf_table* sel = nullptr;
                rq.status = 0;
                try {
                    batch_result(b, (int32_t)i, rq.output_addr, &sel);
                    rq.out_rows = sel ? sel->rows() : 0;
                    if (sel && rq.sel_out && rq.out_rows > 0 && !sel->cols.empty()) {
                        const Column& c0 = sel->cols[0];
                        const size_t nb = std::min<size_t>((size_t)rq.out_rows * c0.type.width(),
                                                           (size_t)rq.sel_cap);
                        RF_HIP(hipMemcpy(rq.sel_out, c0.values, nb, hipMemcpyDeviceToHost));
                    }
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
        // Admit arrived requests into the filling runner.
        const double t = now_us(t0);
        rf_batch* f = runners[filling];
        if (!f->launched) {
            while (next < n && reqs[next].arrival_us <= t) {
                rf_request& rq = reqs[next];
                if (rq.q && rq.q->gj) {
                    {
                        std::lock_guard<std::mutex> lk(wmu);
                        wq.push_back(next);
                    }
                    wcv.notify_one();
                    next++;
                    progress = true;
                    continue;
                }
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
                filling ^= 1;
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
