// HIP runtime layer: device context, device-resident Arrow tables, the
// hiprtc module cache and query execution. Replaces the reference's
// CudaEngine (src/engine/cuda_engine.cpp:25-87) and the host-side
// ArrowTable2 / memman output registry (include/reffine/arrow/table.h,
// src/engine/memory.cpp:5-16).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/reffine_hip.h"
#include "hipgen.hpp"
#include "kernels/rf_abi.h"
#include "lower.hpp"

namespace rf {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArrowError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoMem : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
#define RF_HIP(x) ::rf::hip_check((x), #x)

struct Device {
    int id = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string arch;
    std::mutex mu;
};

// Ahead-of-time kernels (kernels/rf_aot.hip).
hipError_t aot_fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s);
hipError_t aot_index_build(const int64_t* keys, const uint16_t* valid, int64_t n,
                           void* slot_keys, void* slot_rows, uint64_t mask, void* min_key_row,
                           hipStream_t s);
hipError_t aot_index_lookup(const RfProbeDev& h, const int64_t* keys, int64_t n, int64_t* rows,
                            hipStream_t s);
hipError_t aot_copy(const void* src, void* dst, int64_t n16, hipStream_t s);
// Group join kernels (kernels/rf_group.hip).
int64_t aot_sort_hist_entries(int64_t n);
hipError_t aot_radix_sort(void* const keys[2], void* const pay[2][8], const int32_t* pw, int np,
                          int key_bytes, int64_t n, int bits, uint64_t bias, uint32_t* hist,
                          int* which, hipStream_t s);
hipError_t aot_minmax_i64(const int64_t* keys, const int64_t* n_dev, int64_t n_max, void* out,
                          hipStream_t s);
hipError_t aot_adjacent_dups(const int64_t* keys, int64_t n, void* flag, hipStream_t s);
hipError_t aot_not_ascending(const int64_t* keys, int64_t n, void* flag, hipStream_t s);
int64_t aot_scan_entries(int64_t m);
constexpr int64_t kRankFusedWords = int64_t(1) << 18;
hipError_t aot_rank_build_small(const int64_t* keys, int64_t n, const int64_t* n_dev, int64_t base,
                                void* bits, int64_t words, void* cum, void* pack, void* total,
                                const int64_t* k_in, int64_t* k_out, hipStream_t s);
hipError_t aot_rank_place(const int64_t* keys, const int64_t* rows, int64_t n, const int64_t* n_dev,
                          int64_t base, const void* bits, int64_t words, const void* cum,
                          int64_t* keys_out, int64_t* rows_out, hipStream_t s);
hipError_t aot_merge_splits(const int64_t* a, int64_t na, const int64_t* b, int64_t nb,
                            int64_t off_a, int64_t off_b, int64_t tile, int64_t n, int64_t* out,
                            hipStream_t s);
hipError_t aot_scan_u32(void* h, int64_t m, hipStream_t s);
// words x 16 bytes: word w's bits, then its count (RfProbeDev::dense_pack).
hipError_t aot_rank_pack(const void* bits, const void* cum, int64_t words, void* pack,
                         hipStream_t s);
hipError_t aot_rank_rows(const int64_t* keys, int64_t n, int64_t base, const void* bits,
                         const void* cum, int64_t* rows, hipStream_t s);
hipError_t aot_filter_build(const int64_t* keys, int64_t n, void* filter, uint64_t fmask,
                           hipStream_t s);
hipError_t aot_seg_bounds(const uint32_t* rank, int64_t n, int64_t* lo, int64_t* hi,
                          hipStream_t s);

Device& device();          // initialised device (throws if rf_init not called)
Device& init_device(int id);

void* dev_alloc(size_t bytes);  // 256-byte aligned, throws NoMem
void dev_free(void* p);
void dev_fill_u64(void* p, uint64_t v, int64_t n);  // kernel fill on the stream

struct Column {
    std::string format;
    Type type;
    Enc enc = Enc::FLAT;
    int64_t length = 0;       // rows (REE: number of runs)
    int64_t null_count = 0;
    void* values = nullptr;   // device
    void* validity = nullptr; // device bitmap (uint16 words, LSB first) or null
    void* run_ends = nullptr; // REE: device int32 run ends
    int64_t logical_length = 0;
};

// Device allocations of a table, freed when the last holder drops them (the
// table, and any Arrow device array exported from it).
struct OwnedBuffers {
    std::vector<void*> p;
    ~OwnedBuffers();
};

struct HashIndex {
    void* keys = nullptr;
    void* rows = nullptr;
    uint64_t mask = 0;
    int64_t min_key_row = -1;
    // Rank map in place of the slots (keys dense in their range): a bit per
    // key of [dense_base, dense_base + dense_len), the marked keys before
    // each 64-bit word, and the first row of each rank (null: rank == row,
    // the keys being strictly ascending).
    void* dense_bits = nullptr;
    void* dense_cum = nullptr;
    void* dense_rows = nullptr;
    void* dense_pack = nullptr;  // bits and counts interleaved, for probes
    int64_t dense_base = 0;
    uint64_t dense_len = 0;
};

}  // namespace rf

struct rf_table {
    rf_table();
    uint64_t uid;    // unique per table object (a freed table's address may be reused)
    int64_t dim = 1;
    std::vector<rf::Column> cols;
    std::shared_ptr<rf::OwnedBuffers> owned;  // device allocations (shared with exports)
    std::vector<std::shared_ptr<void>> keep;   // borrowed inputs kept alive (imports)
    rf::HashIndex index;
    bool has_index = false;
    // The smallest and largest key of an int64 column 0 without nulls,
    // computed on first use (a table's contents do not change once it
    // exists): 0 not yet computed, 1 known, -1 not applicable or no rows.
    int key_range_state = 0;
    int64_t key_min = 0, key_max = 0;
    rf::Type type() const;
    // Rows of the table: the logical length (a run-end column 0 holds fewer
    // runs than rows).
    int64_t rows() const { return cols.empty() ? 0 : cols[0].logical_length; }
};

namespace rf {

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;      // the entry: rf_q (scan) or rf_gj_fold
    hipFunction_t gather = nullptr;  // rf_q_gather: selection placement (scan modules)
    hipFunction_t bscan = nullptr;   // rf_qb: batched persistent scan
    hipFunction_t bdone = nullptr;   // rf_qb_done: per-unit completion
    hipFunction_t bgather = nullptr; // rf_qb_gather: per-unit selection placement
    hipFunction_t fold_b = nullptr;  // rf_gj_fold_b: folds of several runs (fold modules)
    hipFunction_t fold_w = nullptr;  // rf_gj_fold_w / _wb: a wave per group (large groups)
    hipFunction_t fold_wb = nullptr;
    hipFunction_t agg = nullptr;     // rf_gj_agg: partitioned LDS fold (default mode)
    hipFunction_t lds_done = nullptr; // rf_gj_lds_done: outputs after LDS pre-aggregation
    hipFunction_t emit = nullptr;    // rf_qu_emit: a union merge's rows (entry rf_qu_count)
    hipFunction_t seg_long = nullptr; // rf_seg_fold_long: run folds with long runs
    hipFunction_t seg_wave = nullptr; // rf_seg_fold_wave: a wave per run or window
    int max_blocks_per_cu = 1;
    int bscan_blocks_per_cu = 1;
};

// Compiles (or fetches from the cache) the specialised scan kernel.
Module& jit_module(const KernelSource& ks);
// hiprtc compile for gfx950 only (no device needed); returns the code object.
std::vector<char> jit_compile(const KernelSource& ks);

struct Spec {
    KernelSource ks;
    Module* mod = nullptr;
};

}  // namespace rf

namespace rf {
struct GroupJoinState;
}

struct rf_query {
    rf::Plan plan;
    std::unique_ptr<rf::GroupJoinState> gj;  // group join: sub-queries + workspace
    std::vector<RfProbeDev> probe_override;  // indexes for external probes
    bool det = false;
    rf::Tuning tune;
    std::map<std::string, rf::Spec> specs;  // by nullability signature
    std::map<std::string, rf::Spec> run_specs;  // a run stage's fold, likewise
    // Computed vectors read as tables: run first (on the caller's arguments),
    // their results appended to the arguments (indexed where pre_index).
    std::vector<std::unique_ptr<rf_query>> pre;
    std::vector<bool> pre_index;
    // Plan::composite: one query per component (named by its symbol), run
    // on the same arguments; the output is evaluated over their results.
    std::vector<std::unique_ptr<rf_query>> parts;
    std::vector<std::string> part_names;
    int64_t n_user_args = -1;
    // Workspace (grow-only, reused across runs).
    void* partials = nullptr;
    size_t partials_bytes = 0;
    void* result = nullptr;
    void* counters = nullptr;
    void* wave_arrays = nullptr;  // wave_cnt | wave_dst
    size_t waves_cap = 0;
    void* out_count = nullptr;
    std::vector<void*> outs;      // selection columns in output order
    std::vector<void*> stages;    // per-block staging of the same columns
    size_t out_rows_cap = 0;
    size_t stage_rows_cap = 0;
    int64_t sel_hint = 0;         // selection rows of the last plain run
    void* host_stage = nullptr;  // pinned: result Acc [0, 4096) + count at 4096
    void* host_stage_dev = nullptr;  // the same memory as the device addresses it
    std::mutex mu;
    ~rf_query();
};

namespace rf {

rf_table* table_upload(const ArrowSchema* schema, const ArrowArray* array, int64_t dim);
// Arrow C Device data interface (arrow_device.cpp).
rf_table* table_import_device(const ArrowSchema* schema, ArrowDeviceArray* arr, int64_t dim);
void table_export_device(const rf_table* t, ArrowSchema* schema, ArrowDeviceArray* arr);
bool format_type_of(const std::string& f, Type& t);
std::string arrow_format_of(const Type& t);
rf_table* table_wrap(int64_t dim, int64_t n_cols, int64_t n_rows, const char* const* formats,
                     void* const* values, void* const* validity);
void table_build_index(rf_table* t);
void table_free(rf_table* t);
RfProbeDev probe_of(const rf_table* t);

struct Launch {
    Spec* spec = nullptr;
    RfScanArgs args{};
    int grid = 0;
    bool empty = false;
    int64_t lo = 0, end = 0;
};

// Prepares arguments and workspace for one run (no launch).
Launch prepare(rf_query* q, const rf_arg* args, int64_t n_args, bool* ok);
// The compiled scan kernel of `q` for the arguments' null layout, or null
// when none was compiled for it (nothing is compiled here).
Spec* find_spec(rf_query* q, const rf_arg* args, int64_t n_args);
// Blocks of a scan launch: waves_per_cu resident waves per CU (bounded by
// occupancy), no more than the wave-tiles need.
int scan_grid(const Launch& l);
// The no-allocation run (noalloc.hip): per-device scratch set up once, with
// the device current, by init_device; then rf_query_run_noalloc's body.
void noalloc_setup(int device);
void run_query_noalloc(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
                       void* const* out_cols, int64_t out_cap, int64_t* out_rows);
// Arguments of one run (columns, probes, constants, rows, tiles) without
// the query's workspace pointers.
Launch prepare_args(rf_query* q, const rf_arg* args, int64_t n_args);
// Host-side finish of a run: the function output from the reductions'
// device result bytes `acc` and the selection table `sel`.
void finish_output(rf_query* q, const KernelSource& ks, const rf_arg* args, int64_t n_args,
                   const char* acc, bool empty, rf_table* sel, void* output_addr,
                   rf_table** out_table);
// Checks argument kinds and table schemas against the plan's inputs.
void check_args(const Plan& plan, const rf_arg* args, int64_t n_args);
// A new device table of `count` rows (owned columns, no validity).
rf_table* alloc_table(const std::vector<Type>& types, int64_t count, int64_t dim);
// A table over borrowed device columns (freed with table_free, which then
// releases only the host descriptor).
rf_table* table_view(const std::vector<Type>& types, const std::vector<void*>& values,
                     int64_t count, int64_t dim);
void launch(rf_query* q, Launch& l, hipStream_t s, bool scan_only);
// Launch + download + host-side finish; returns the output.
void run_query(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
               rf_table** out_table);
// Mean ms per run of all the query's kernels, and of its scan kernel alone.
void time_query(rf_query* q, const rf_arg* args, int64_t n_args, int iters, double* ms_all,
                double* ms_scan);

// Output struct layout (natural alignment, ScalarPass field order).
void c_layout(const Type& t, size_t& size, size_t& align);

// Group join: sub-queries for the build and probe scans, the per-group
// fold module and a grow-only workspace.
// The resident build side of a group join for one build table and one set
// of build-plan scalar arguments: the qualifying keys sorted (a key's rank
// is its position), their build rows, the key -> rank index and its bitmap
// filter. Shared (reference-counted) between plain runs and batch rounds.
struct GjBuild {
    int64_t groups = 0;     // qualifying build keys (0: the join is empty)
    int64_t kmin = 0;
    void* keys = nullptr;   // i64 [groups], ascending
    void* rows = nullptr;   // i64 [groups], build row of each key
    void* hkeys = nullptr;  // index key -> rank
    void* hrows = nullptr;
    uint64_t index_cap = 0;
    void* filter = nullptr;
    uint64_t filter_bits = 0;
    // Rank map over [kmin, kmin + dense_len) for dense key ranges (then no
    // index or filter): key bits and the words' first ranks.
    void* dense_bits = nullptr;
    void* dense_cum = nullptr;
    void* dense_pack = nullptr;  // bits and counts interleaved, for probes
    uint64_t dense_len = 0;
    // Duplicate-key check not yet read: the count of marked keys, written by
    // the build into pinned host memory, must equal `groups` once the
    // stream has been synchronised (then the build is usable).
    const volatile uint64_t* pending_total = nullptr;
    // A build launched without a host synchronisation: the count of
    // qualifying keys is still on the device (groups_dev, at most
    // groups_ub); the build's last kernel copies it to pending_groups
    // (pinned memory), which gj_verify reads into `groups`.
    int64_t groups_ub = 0;
    const int64_t* groups_dev = nullptr;
    const volatile int64_t* pending_groups = nullptr;
    RfProbeDev probe() const;  // the probe plan's index for the external probes
    ~GjBuild();
};

struct GroupJoinState {
    std::unique_ptr<rf_query> build_q, probe_q;
    // The probe scan against a hash index (sparse build keys) prefers
    // another launch shape than against a direct map: compiled on first use
    // unless the caller fixed the tuning.
    std::unique_ptr<rf_query> probe_q_hash;
    // The probe scans that pre-aggregate the joined rows in LDS (default
    // mode, few groups), against a rank map and against a hash index:
    // compiled on first use.
    std::unique_ptr<rf_query> probe_q_lds, probe_q_lds_hash;
    bool probe_tuned = false;
    std::map<std::string, Spec> fold_specs;  // by nullability signature
    int table_scale = 2;                     // index capacity >= scale * groups
    void* small = nullptr;                   // key range, duplicate flag, scratch
    void* keys2 = nullptr;                   // sort ping-pong buffers (build side)
    void* rows2 = nullptr;
    size_t keys_cap = 0;
    void* hist = nullptr;
    size_t hist_cap = 0;
    void* rank2 = nullptr;                   // sort ping-pong buffers (joined rows)
    std::vector<void*> pay2;
    size_t join_cap = 0;
    void* seg = nullptr;                     // lo | hi per group
    size_t seg_cap = 0;
    // Partitioned aggregation (default mode): zero partition starts for an
    // empty probe, and the partitioned joined rows (sized to the probe
    // scan's staging; the block counts go in `hist`).
    void* part_small = nullptr;
    void* prank = nullptr;
    std::vector<void*> ppay;
    size_t part_cap = 0;
    void* lds_part = nullptr;                // pre-aggregation: the blocks' LDS slots
    size_t lds_part_cap = 0;
    // Build sides kept resident, by build key (the build table's uid and the
    // build plan's scalar arguments); the oldest is dropped beyond kBuilds.
    static constexpr size_t kBuilds = 16;
    std::map<std::string, std::shared_ptr<GjBuild>> builds;
    std::vector<std::string> build_order;
    ~GroupJoinState();
};

// The build side of group join `q` for these arguments: from the cache, or
// built on the library stream (synchronous). Caller holds q->mu.
// defer_check: a dense build's duplicate-key check is left for the caller's
// own synchronisation (gj_verify after it) instead of a synchronisation here.
std::shared_ptr<GjBuild> gj_build(rf_query* q, const rf_arg* args, int64_t n_args,
                                  bool defer_check = false);
// After the stream is synchronised, with q->mu held: throws (and drops the
// cached builds) when a deferred duplicate-key check failed.
void gj_verify(rf_query* q, GjBuild& b);
// The probe sub-query for a build side (direct map or hash index).
rf_query* gj_probe_query(rf_query* q, const GjBuild& b);
// Drops the cached build sides (in-flight batch rounds keep theirs).
void gj_clear_cache(rf_query* q);
// The fold module of group join `q` for the arguments' nullability.
Spec& gj_fold_spec(rf_query* q, const rf_arg* args, int64_t n_args);
// Fold arguments of one group join run, apart from the joined rows (pay,
// seg_lo, seg_hi) and the output columns.
void gj_fold_args(rf_query* q, const Spec& fs, const GjBuild& b, const rf_arg* args,
                  int64_t n_args, RfFoldArgs& A);
int bits_of(uint64_t range);
// Joined rows per group (on average) from which a wave folds each group.
constexpr int64_t kWaveFoldRows = 64;
int64_t aot_rank_cum_entries(int64_t words);
hipError_t aot_rank_scatter(const int64_t* keys, const int64_t* rows, int64_t n, int64_t base,
                            const void* bits, int64_t words, const void* cum, int64_t* keys_out,
                            int64_t* rows_out, void* total, hipStream_t s);
hipError_t aot_rank_build(const int64_t* keys, int64_t n, int64_t base, void* bits, int64_t words,
                          void* cum, hipStream_t s);
hipError_t aot_gj_concat(const RfCatSeg* segs, int32_t nseg, int64_t max_rows,
                         const RfCatArgs& A, hipStream_t s);
hipError_t aot_partition(const RfPartArgs& P, hipStream_t s);
int64_t aot_part_entries(int64_t nparts, int64_t nwaves, int64_t waves_per_block);

// Batched execution (engine_batch.cpp).
rf_batch* batch_create(int cap, int64_t arena_bytes, bool use_graph);
int64_t batch_footprint(rf_query* q, const rf_arg* args, int64_t n_args);
bool batch_add(rf_batch* b, rf_query* q, const rf_arg* args, int64_t n_args, int32_t* unit);
void batch_launch(rf_batch* b);
bool batch_poll(rf_batch* b);
void batch_wait(rf_batch* b);
void batch_result(rf_batch* b, int32_t unit, void* output_addr, rf_table** out_table);
void batch_reset(rf_batch* b);
int batch_units(const rf_batch* b);
double serve(rf_request* reqs, int64_t n, int32_t max_units, int64_t arena_bytes, bool use_graph,
             int32_t runners);
double serve_plain(rf_request* reqs, int64_t n);

// Creates the sub-queries of a group join and compiles its all-valid
// kernels; `probe_tune` (optional) sets the probe scan's launch shape.
void setup_group_join(rf_query* q, const Tuning* probe_tune);
// Kernel sources of a group join for a nullability (compile checks).
std::vector<KernelSource> group_join_sources(const Plan& plan,
                                             const std::vector<std::vector<bool>>& nullable);

}  // namespace rf
