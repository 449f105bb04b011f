/*
 * reffine_hip.h — C ABI of the MI355X (gfx950) execution path for reffine
 * queries.
 *
 * Every function returns an int status (RF_OK == 0). On failure the
 * calling thread's last error message is available from rf_last_error().
 * Nothing in this library calls exit() or abort() across the ABI.
 *
 * The reference launches GPU code through CudaEngine
 * (include/reffine/engine/cuda_engine.h:24-35, src/engine/cuda_engine.cpp:
 * 25-87) and runs queries compiled by compile_op<T>/compile_loop<T>
 * (include/reffine/utils/utils.h:20-54), whose generated entry point is
 * `extern "C" void <name>(OutT* output_addr, In0 a0, ...)`
 * (src/pass/canonpass.cpp:50-64, src/pass/cemitter.cpp:227-252).
 * Each declaration below names the reference interface it replaces.
 */
#ifndef REFFINE_HIP_H_
#define REFFINE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arrow C data interface (identical to include/reffine/arrow/abi.h:18-49). */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
    const char* format;
    const char* name;
    const char* metadata;
    int64_t flags;
    int64_t n_children;
    struct ArrowSchema** children;
    struct ArrowSchema* dictionary;
    void (*release)(struct ArrowSchema*);
    void* private_data;
};
struct ArrowArray {
    int64_t length;
    int64_t null_count;
    int64_t offset;
    int64_t n_buffers;
    int64_t n_children;
    const void** buffers;
    struct ArrowArray** children;
    struct ArrowArray* dictionary;
    void (*release)(struct ArrowArray*);
    void* private_data;
};
#endif /* ARROW_C_DATA_INTERFACE */

/* Arrow C Device data interface (arrow/c/abi.h): an ArrowArray with the
 * device that holds its buffers. */
#ifndef ARROW_C_DEVICE_DATA_INTERFACE
#define ARROW_C_DEVICE_DATA_INTERFACE
typedef int32_t ArrowDeviceType;
#define ARROW_DEVICE_CPU 1
#define ARROW_DEVICE_CUDA 2
#define ARROW_DEVICE_CUDA_HOST 3
#define ARROW_DEVICE_ROCM 10
#define ARROW_DEVICE_ROCM_HOST 11
struct ArrowDeviceArray {
    struct ArrowArray array;
    int64_t device_id;
    ArrowDeviceType device_type;
    void* sync_event;
    int64_t reserved[3];
};
#endif /* ARROW_C_DEVICE_DATA_INTERFACE */

enum rf_status {
    RF_OK = 0,
    RF_ERR_INVALID = 1,     /* bad argument, malformed IR, type error      */
    RF_ERR_UNSUPPORTED = 2, /* IR shape the GPU lowering does not accept   */
    RF_ERR_HIP = 3,         /* HIP runtime / hiprtc failure                */
    RF_ERR_NOMEM = 4,       /* device memory or admission limit exceeded   */
    RF_ERR_ARROW = 5,       /* Arrow input the device layer cannot take    */
    RF_ERR_BUSY = 6         /* server queue full / no free slot            */
};

/* Opaque handles. */
typedef struct rf_table rf_table;   /* device-resident Arrow record batch   */
typedef struct rf_query rf_query;   /* lowered + JIT-specialised query plan */

/* One argument of a compiled query, in Func::inputs order. A VECTOR input
 * is a device table; a primitive input is passed by value (int64 for all
 * integer / idx / bool types, double for float types), matching the by-value
 * primitives of the generated entry point (include/reffine/base/type.h:
 * 212-249). */
typedef struct rf_arg {
    int32_t kind; /* 0 = table, 1 = int64, 2 = double */
    int32_t pad;
    union {
        rf_table* table;
        int64_t i64;
        double f64;
    } u;
} rf_arg;

/* Result of a query. A reduction writes its state struct (C layout, fields
 * in declaration order, natural alignment — the ScalarPass output layout,
 * src/pass/scalarpass.cpp:103-227) to `output_addr`. An Op writes a device
 * table handle into `*out_table` (the reference writes an ArrowTable* into
 * its ArrowTable** output, src/pass/loopgen.cpp:113-204). */

/* ---------------------------------------------------------------- device */

/* Replaces CudaEngine::Get() (src/engine/cuda_engine.cpp:25-43): selects
 * the HIP device and creates the library's stream. Idempotent. */
int rf_init(int device);

/* Thread-local message of the last failing call ("" when none). */
const char* rf_last_error(void);

/* Device properties: CU count, HBM bytes, free bytes. */
int rf_device_info(int32_t* num_cus, int64_t* total_bytes, int64_t* free_bytes);

/* Blocks until all work issued by the library on its stream is finished
 * (replaces cuCtxSynchronize, main.cpp:229). */
int rf_sync(void);

/* ----------------------------------------------------------------- tables */

/* Uploads a host Arrow record batch (struct array "+s", one child per
 * column: formats c,C,s,S,i,I,l,L,f,g,b? and "+r" run-end children) into
 * HBM. Replaces the host ArrowTable2 built by load_arrow_file
 * (src/utils/utils.cpp:11-34, include/reffine/arrow/table.h:24-60): value
 * buffers and validity bitmaps are copied to 256-byte aligned device
 * memory. `dim` is the table's iterator dimension (reffine's
 * ArrowTable::dim, include/reffine/arrow/abi.h:56-69). Offsets are
 * honoured (the reference's accessors ignore them, vinstr.h:25-111): a
 * sliced batch -- offset on the struct or on its children -- uploads only
 * its rows; validity bits at any bit offset; a run-end column keeps the
 * runs covering the slice, cut to it. */
int rf_table_upload(const struct ArrowSchema* schema,
                    const struct ArrowArray* array, int64_t dim,
                    rf_table** out);

/* Imports a record batch given through the Arrow C Device data interface
 * (struct array "+s"). The array is moved in (its release callback runs
 * here after a copy or on error, or when the table is freed). ROCm memory on
 * this library's device is borrowed without a copy when every column is flat
 * with 16-byte aligned values and validity on a 16-bit word boundary (the
 * kernels' load granules), else copied; host memory (CPU, pinned) is
 * uploaded as rf_table_upload does. A ROCm array's sync_event (hipEvent_t*)
 * is waited on before use. The schema is read, not released. Replaces the
 * host-only ArrowTable exchange (include/reffine/arrow/abi.h:56-69,
 * src/utils/utils.cpp:11-34). */
int rf_table_import_device(const struct ArrowSchema* schema, struct ArrowDeviceArray* array,
                           int64_t dim, rf_table** out);

/* Exports a device table as an Arrow C Device array (device_type ROCm, no
 * copy, sync_event NULL: the data is complete). The exported structs share
 * the table's device buffers and keep them alive after rf_table_free until
 * their release callbacks run; the caller owns both structs. A table over
 * borrowed buffers (rf_table_wrap_device, batch results) exports views that
 * must not outlive the buffers' owner. */
int rf_table_export_device(const rf_table* t, struct ArrowSchema* schema,
                           struct ArrowDeviceArray* array);

/* Creates a device table over column buffers that are already in device
 * memory (borrowed; the caller keeps them alive). `formats[i]` is the Arrow
 * format string of column i; `validity[i]` may be NULL (all valid). */
int rf_table_wrap_device(int64_t dim, int64_t n_cols, int64_t n_rows,
                         const char* const* formats, void* const* values,
                         void* const* validity, rf_table** out);

/* Number of rows, columns, and a column's device value / validity pointers.
 * Replaces the vinstr accessors get_vector_len / get_vector_data_buf /
 * get_vector_bit_buf (include/reffine/vinstr/vinstr.h:43-69). */
int rf_table_shape(const rf_table* t, int64_t* n_rows, int64_t* n_cols);
int rf_table_column(const rf_table* t, int64_t col, void** values,
                    void** validity, int64_t* length, char* format_out,
                    size_t format_cap);

/* Copies column `col` values (and, if `validity_out` is non-NULL, its
 * validity bitmap, (length+7)/8 bytes, all-ones when the column has none)
 * to host memory. */
int rf_table_download_column(const rf_table* t, int64_t col, void* values_out,
                             size_t values_cap, uint8_t* validity_out,
                             size_t validity_cap);

/* Builds the device index over int64 column 0 (key -> first valid row; for
 * a run-end column 0, key -> run, as the reference indexes the run values): a
 * rank map (a bit per key of the range, counts per 64 keys, first rows) when
 * the column has no nulls and its keys span at most 32 values per row plus
 * 2^20, open addressing otherwise. Replaces ArrowTable2::build_index /
 * build_vector_index (include/reffine/arrow/table.h:107-138,
 * src/vinstr/external.cpp:11-16). */
int rf_table_build_index(rf_table* t);

/* Looks up `n` keys in the index built above; writes the row or -1
 * (vector_locate, include/reffine/vinstr/vinstr.h:137-146). Device
 * pointers. */
int rf_table_locate(const rf_table* t, const int64_t* d_keys, int64_t n,
                    int64_t* d_rows_out);

int rf_table_free(rf_table* t);

/* ---------------------------------------------------------------- queries */

/* Compiles a reffine Func given as IR text (the S-expression form written
 * by reffine_amd.reffiner, one node per reffine IR node). Runs the GPU
 * lowering (the counterpart of LoopGen, src/pass/loopgen.cpp:44-269) and
 * specialises the hand-written kernels for the query's row expressions.
 * Replaces compile_op<T>(func, vectorize) (include/reffine/utils/utils.h:
 * 47-54). `flags`: RF_COMPILE_DETERMINISTIC selects the fixed-order
 * reduction mode; RF_COMPILE_INDEX_SCALE(s), s in 2..15, sizes a group
 * join's hash index over sparse build keys at >= s slots per key (the
 * load factor the autotuner picks; default 2; a rank map over dense keys
 * does not use it). */
#define RF_COMPILE_DETERMINISTIC 1u
#define RF_COMPILE_INDEX_SCALE(s) ((uint32_t)(s) << 8)
int rf_compile(const char* ir_text, uint32_t flags, rf_query** out);

/* rf_compile with an explicit launch shape `tuning` = {BLOCK (64..1024,
 * multiple of 64), ITEMS (1, 2, 4, 8 or 16), SUB (1..64), WAVES_PER_CU
 * (1..64, 0 = 16), PIPELINE (0/1: overlap each wave's next load group with
 * the current one), LSUB (sub-steps per load group, dividing SUB; 0 = from
 * the register budget)}: the autotuner's entry point. A wave-tile is
 * SUB*64*ITEMS rows, loaded LSUB*64*ITEMS rows at a time. */
int rf_compile_tuned(const char* ir_text, uint32_t flags,
                     const int32_t* tuning, rf_query** out);

/* Lowers and JIT-compiles a Func for gfx950 without a device (no module
 * load): the build-time check that a query shape is accepted. `flags` bit 0
 * as rf_compile; bit 1 compiles the variant with validity bitmaps on every
 * column. Writes the lowering summary and kernel source to `buf`. No
 * reference counterpart (the reference compiles only inside compile_op). */
int rf_compile_check(const char* ir_text, uint32_t flags, char* buf,
                     size_t cap);

/* Byte size of the query's reduction output struct (0 for Op outputs), and
 * whether the output is a table. */
int rf_query_output_info(const rf_query* q, int64_t* out_bytes,
                         int32_t* out_is_table);

/* Launch shape of the query's scan kernel, written to tuning[0..5]:
 * threads per block, rows per lane per sub-step (ITEMS), sub-steps per
 * wave-tile (SUB), resident waves per CU, the pipelining flag and LSUB
 * (0 = from the register budget). In the deterministic mode ITEMS and SUB
 * fix the reduction order (wave-tiles of SUB*64*ITEMS rows, the completion
 * combine over BLOCK threads). No reference counterpart. */
int rf_query_tuning(const rf_query* q, int32_t* tuning);

/* Human-readable lowering summary (kernel kind, columns, fields). */
int rf_query_describe(const rf_query* q, char* buf, size_t cap);

/* Drops the resident build sides a group join query keeps (per build table
 * and build scalars); the next run builds again. A no-op for other queries.
 * No reference counterpart (the reference evaluates the nested loop from
 * scratch on every call, src/pass/loopgen.cpp:113-269). */
int rf_query_clear_cache(rf_query* q);

/* Replaces the call of the generated entry point `void name(OutT*, ...)`:
 * runs the query on device tables and synchronises. `output_addr` receives
 * the reduction struct (host memory, rf_query_output_info bytes); for Op
 * queries `out_table` receives a new device table (free with
 * rf_table_free). */
int rf_query_run(rf_query* q, const rf_arg* args, int64_t n_args,
                 void* output_addr, rf_table** out_table);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): runs a scan query (reductions and at most one selection
 * over one table, probes of indexed tables; not group joins, run folds or
 * computed vectors) as rf_query_run does, with no hipMalloc, hipFree or
 * other allocating call. The kernels' scratch (partials, counter, staging)
 * lives in one fixed device array of RF_NOALLOC_SCRATCH_BYTES per GPU; a run
 * that needs more fails with RF_ERR_NOMEM before launching. `output_addr`
 * receives the reduction struct (host memory; a field holding the selection
 * table gets NULL). The selection's columns (rf_query_output_info order:
 * iterators then outputs) are placed into the device buffers `out_cols[j]`,
 * each with room for `out_cap` rows; `*out_rows` receives the row count
 * (RF_ERR_NOMEM when it exceeds out_cap; the first out_cap rows are placed).
 * The kernel for the tables' null layout and index forms must be loaded
 * (rf_compile loads the no-null one for probed tables whose index is not a
 * packed rank map over ascending keys; a run through rf_query_run loads the
 * others) and probed tables indexed, else RF_ERR_INVALID. Calls share the device's scratch and
 * synchronise before returning; calls that overlap on one GPU race. */
#define RF_NOALLOC_SCRATCH_BYTES (64ll << 20)
int rf_query_run_noalloc(rf_query* q, const rf_arg* args, int64_t n_args,
                         void* output_addr, void* const* out_cols, int64_t out_cap,
                         int64_t* out_rows);

/* Times `iters` back-to-back runs of the query's kernels with HIP events on
 * the library stream (no host work): `ms_per_run` is the mean device time of
 * a whole run (scan + selection gather), `ms_scan_kernel` (optional) that of
 * the kernel that streams the input table alone: the scan kernel, or for a
 * function with reductions over the runs of a run-end table, the run fold
 * (whose runs table the scan then reads). */
int rf_query_time(rf_query* q, const rf_arg* args, int64_t n_args,
                  int32_t iters, double* ms_per_run, double* ms_scan_kernel);

int rf_query_free(rf_query* q);

/* ------------------------------------------------------ batched / served */

/* A batch runner: a round of work units (one compiled query over its
 * arguments each) run together -- per distinct query kernel, one persistent
 * launch over the list of all its units' (unit, wave-tile) items (each wave
 * a contiguous range), then per-unit completion and selection placement.
 * A group join unit runs its probe scan in the round; after the round one
 * sort and one fold launch serve all its group join units. Each unit's
 * result equals the deterministic-mode run of the same query alone
 * (RF_COMPILE_DETERMINISTIC queries give identical bits in both), with one
 * exception: a unit whose function folds the runs of a run-end table or
 * reads a computed vector as a table runs on the plain path, synchronously
 * inside rf_batch_add (under rf_serve: at admission, while the serving loop
 * waits for it), with the query's own flags, so its result is that of
 * rf_query_run. `max_units` units per round;
 * `arena_bytes` of device memory for the units' partials and selection
 * buffers (admission by footprint: rf_batch_add returns RF_ERR_BUSY when a
 * unit does not fit); `use_graph` != 0 captures a round into a hipGraph once
 * per set of query kernels and replays it. No reference counterpart (the
 * reference runs one query per call, benchmark/main.cpp:76-82). */
typedef struct rf_batch rf_batch;
int rf_batch_create(int32_t max_units, int64_t arena_bytes, int32_t use_graph,
                    rf_batch** out);
/* Arena bytes one unit of `q` over `args` needs. */
int rf_batch_footprint(rf_query* q, const rf_arg* args, int64_t n_args,
                       int64_t* bytes);
/* Adds a unit (the arguments are copied; the tables must stay alive until
 * the round is reset). RF_ERR_BUSY: the round is full. */
int rf_batch_add(rf_batch* b, rf_query* q, const rf_arg* args, int64_t n_args,
                 int32_t* unit);
/* Launches the round asynchronously on the runner's own stream. */
int rf_batch_launch(rf_batch* b);
/* *done = 1 once the launched round has finished (no wait; launches the
 * group join sort and fold once the round's scans are done). */
int rf_batch_poll(rf_batch* b, int32_t* done);
int rf_batch_wait(rf_batch* b);
/* Output of a unit of a finished round, as rf_query_run gives it; a
 * selection (or group join) table is a view of the runner's arena, valid until
 * rf_batch_reset (free it with rf_table_free). A unit whose output is a
 * table needs `out_table` (RF_ERR_INVALID without it). */
int rf_batch_result(rf_batch* b, int32_t unit, void* output_addr,
                    rf_table** out_table);
/* Waits for the round if needed and starts an empty one. */
int rf_batch_reset(rf_batch* b);
int rf_batch_free(rf_batch* b);

/* One request of the serving loop. Inputs: the compiled query, its
 * arguments, the arrival time (microseconds from the start of the run,
 * ascending over the array), where to write the reduction output
 * (output_addr, may be NULL) and the selection's first column (sel_out,
 * sel_cap bytes, may be NULL). Outputs: out_rows (selection rows),
 * done_us (completion time) and status (RF_OK or an error code). */
typedef struct rf_request {
    rf_query* q;
    const rf_arg* args;
    int64_t n_args;
    double arrival_us;
    void* output_addr;
    void* sel_out;
    int64_t sel_cap;
    int64_t out_rows;
    double done_us;
    int32_t status;
    int32_t pad;
} rf_request;

/* Serves `n` requests as they arrive (a real-time loop against the arrival
 * times) with continuous batching over `runners` batch runners (2..4; 0
 * selects 3), each on its own stream: requests that arrive while rounds are
 * on the GPU are admitted into an idle runner as units and arena memory
 * allow, and it launches at once, so up to `runners` rounds overlap on the
 * device. Latency of request i = done_us - arrival_us. `wall_us` receives
 * the time from start to the last completion. The runners (arena, buffers,
 * captured graphs) are kept for the next call with the same max_units,
 * arena_bytes and use_graph (at most four runners in all, held until the
 * process ends). No reference counterpart. */
int rf_serve(rf_request* reqs, int64_t n, int32_t max_units, int64_t arena_bytes,
             int32_t use_graph, int32_t runners, double* wall_us);

/* The plain counterpart of rf_serve for the same requests: each one, once
 * it has arrived, runs alone through rf_query_run (its own launches and
 * host synchronisation), in arrival order. */
int rf_serve_plain(rf_request* reqs, int64_t n, double* wall_us);

/* Times a plain 16-bytes-per-lane device copy of `bytes` (read + write)
 * over `iters` launches; the measured HBM ceiling the scan kernels are
 * compared with, and the reference size for calibrating byte counters. No
 * reference counterpart. */
int rf_calib_copy(int64_t bytes, int32_t iters, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* REFFINE_HIP_H_ */
