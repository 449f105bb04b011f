// The no-allocation run of a scan query (rf_query_run_noalloc).
//
// rf_query_run keeps a grow-only workspace per query and allocates the
// returned selection table; callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) call this variant instead. It runs the same kernels
// (rf_q + rf_q_gather of the query's module) with their scratch -- the
// completion counter, the result struct and selection count, the per-wave
// partials, per-wave counts and offsets, and the selection staging -- laid
// out in one fixed device array per GPU, and places the selection straight
// into column buffers the caller passes. No hipMalloc, hipFree or other
// allocating routine is called.
//
// The array is `scratch` below: 64 MiB, held by every GPU the library runs
// on until the process ends. Its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current (the
// library's init); host code uses that address, never the array's name.
// Every call on a GPU shares the array, which keeps the last call's
// contents: the counter is cleared by a fill kernel on each call (the other
// regions are written before they are read), and calls that overlap on one
// GPU race.
#include <hip/hip_runtime.h>

#include <cstring>

#include "engine.hpp"

namespace rf {

constexpr size_t kScratchBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];  // the array's address on each device

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw HipError("device id out of range");
    if (g_scratch[device]) return;
    void* p = nullptr;
    RF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[device] = p;
}

static size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

void run_query_noalloc(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
                       void* const* out_cols, int64_t out_cap, int64_t* out_rows) {
    if (!q->pre.empty() || q->plan.runs || q->plan.seg || q->gj ||
        q->plan.source == Plan::MERGE || q->plan.composite)
        throw Unsupported("the no-allocation run takes scan queries only (reductions and one "
                          "selection over one table): not group joins, run folds, union merges "
                          "or computed vectors");
    Device& d = device();
    char* base = (char*)g_scratch[d.id];
    if (!base) throw HipError("the scratch array of this device is not set up");
    if (out_cap < 0) throw IRError("out_cap is negative");
    std::lock_guard<std::mutex> lk(q->mu);
    check_args(q->plan, args, n_args);
    // Nothing is compiled or built here: the kernel for this null layout and
    // the probed tables' indexes must exist.
    for (const Probe& pr : q->plan.probes) {
        if (pr.external) throw Unsupported("external probes need a group join's build side");
        if (!args[pr.table].u.table->has_index)
            throw IRError("a probed table has no index: call rf_table_build_index first");
    }
    if (!find_spec(q, args, n_args))
        throw IRError("no kernel is loaded for these tables' null layout and index forms: run "
                      "the query once through rf_query_run on tables with the same validity "
                      "buffers and index forms (packed rank map over ascending keys, or not)");
    Launch l = prepare_args(q, args, n_args);
    const KernelSource& ks = l.spec->ks;
    if ((size_t)ks.acc_bytes > 4096) throw Unsupported("reduction state larger than 4 KiB");
    RfScanArgs& A = l.args;
    l.grid = scan_grid(l);
    const int64_t nw = (int64_t)l.grid * (ks.tune.block / 64);
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;

    // Scratch layout: counter | result (4 KiB) + count | partials | wave
    // counts and offsets | staging columns.
    size_t off = 0;
    const size_t o_cnt = off;
    off += 256;
    const size_t o_res = off;
    off += 4096 + 256;
    const size_t nparts = ks.det ? (size_t)std::max<int64_t>(A.ntiles, 1) : (size_t)nw;
    const size_t o_part = off;
    off += up256(nparts * std::max<size_t>(ks.acc_bytes, 8));
    size_t o_waves = 0;
    std::vector<size_t> o_stage;
    int64_t stage_rows = 0;
    if (ks.has_sel) {
        stage_rows = ((A.ntiles + nw - 1) / nw) * tile;
        const size_t stage_total = (size_t)std::max<int64_t>(stage_rows * nw, 1);
        o_waves = off;
        off += up256((size_t)nw * 16);
        for (auto& t : ks.stage_types) {
            o_stage.push_back(off);
            off += up256(stage_total * t.width());
        }
        if (!out_cols)
            throw IRError("a selection needs out_cols (one device buffer per output column)");
        for (size_t j = 0; j < ks.out_types.size(); j++)
            if (!out_cols[j]) throw IRError("out_cols[" + std::to_string(j) + "] is NULL");
    }
    if (off > kScratchBytes)
        throw NoMem("the run needs " + std::to_string(off) + " bytes of scratch; the "
                    "no-allocation run has " + std::to_string(kScratchBytes) +
                    " (fewer rows, or rf_query_run)");

    A.counters = (uint32_t*)(base + o_cnt);
    A.result = base + o_res;
    A.out_count = (int64_t*)(base + o_res + 4096);
    A.partials = base + o_part;
    if (ks.has_sel) {
        A.stage_rows = stage_rows;
        A.nwaves = nw;
        A.wave_cnt = (int64_t*)(base + o_waves);
        A.wave_dst = A.wave_cnt + nw;
        for (size_t j = 0; j < o_stage.size(); j++) A.stage[j] = base + o_stage[j];
        for (size_t j = 0; j < ks.out_types.size(); j++) A.out[j] = out_cols[j];
        A.out_cap = out_cap;
    }
    char host[4096 + 8];
    std::memset(host, 0, sizeof host);
    if (!l.empty) {
        dev_fill_u64(A.counters, 0, 32);  // the completion counter starts at zero
        launch(q, l, d.stream, false);
        RF_HIP(hipMemcpyAsync(host, A.result, sizeof host, hipMemcpyDeviceToHost, d.stream));
    }
    RF_HIP(hipStreamSynchronize(d.stream));
    const int64_t count = (l.empty || !ks.has_sel) ? 0 : *(const int64_t*)(host + 4096);
    if (out_rows) *out_rows = count;
    if (count > out_cap)
        throw NoMem("the selection has " + std::to_string(count) + " rows; out_cap is " +
                    std::to_string(out_cap) + " (the first out_cap rows were placed)");
    // A table output's rows are in out_cols; a struct output holding the
    // selection gets a NULL table handle in that field.
    if (!q->plan.output_type.is_vector())
        finish_output(q, ks, args, n_args, host, l.empty, nullptr, output_addr, nullptr);
}

}  // namespace rf
