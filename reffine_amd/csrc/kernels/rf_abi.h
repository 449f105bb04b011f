// Launch ABI shared by the host runtime and the device skeleton.
// Compiled by the host compiler, by hipcc (ahead-of-time kernels) and by
// hiprtc (query-specialised kernels); layouts must not depend on which.
#pragma once

#ifdef __HIPCC_RTC__
typedef __INT8_TYPE__ int8_t;
typedef __INT16_TYPE__ int16_t;
typedef __INT32_TYPE__ int32_t;
typedef __INT64_TYPE__ int64_t;
typedef __UINT8_TYPE__ uint8_t;
typedef __UINT16_TYPE__ uint16_t;
typedef __UINT32_TYPE__ uint32_t;
typedef __UINT64_TYPE__ uint64_t;
#else
#include <stdint.h>
#endif

#define RF_MAXCOLS 32
#define RF_MAXPARAMS 96
#define RF_MAXPROBES 4
#define RF_MAXOUT 16
// A union merge (rf_qu_count / rf_qu_emit): RF_MERGE_TILE merge positions
// per block of RF_MERGE_BLOCK threads, RF_MERGE_IT per thread.
#define RF_MERGE_IT 4
#define RF_MERGE_BLOCK 256
#define RF_MERGE_TILE (RF_MERGE_IT * RF_MERGE_BLOCK)

#define RF_EMPTY_KEY ((int64_t)0x8000000000000000ull)

union RfParam {
    int64_t i;
    double f;
};

// Open-addressing table over an int64 key column: first valid row per key
// (the reference's build_index keeps the first occurrence,
// include/reffine/arrow/table.h:107-138).
struct RfProbeDev {
    const int64_t* keys;   // slot keys, RF_EMPTY_KEY marks a free slot
    const int64_t* rows;   // slot row (minimum row holding the key)
    uint64_t mask;         // capacity - 1 (capacity a power of two)
    int64_t min_key_row;   // row of key == RF_EMPTY_KEY, -1 if none
    const uint32_t* filter;  // optional bitmap of (hash >> 32) & fmask over the keys
    uint64_t fmask;          // filter bits - 1 (a power of two)
    // Optional rank map for a dense key range (replaces the slots): bit
    // k - dense_base of dense_bits marks key k, for k in [dense_base,
    // dense_base + dense_len); its row is the number of marked keys below
    // it: dense_cum[word] (marked keys in earlier 64-bit words) plus the
    // marked bits below it in its word. Keys outside the range have no row.
    const uint64_t* dense_bits;
    const uint32_t* dense_cum;
    int64_t dense_base;
    uint64_t dense_len;
    const int64_t* dense_rows;  // row of each rank; null: the rank is the row
    // The same map packed for probes: 16 bytes per 64-bit word, the word's
    // key bits then (low 32 bits) the marked keys before it -- one load per
    // probe. Probes use it when it is set.
    const uint64_t* dense_pack;
};

struct RfScanArgs {
    int64_t start;        // first row of tile 0 (lo rounded down to ITEMS)
    int64_t lo, end;      // driver rows [lo, end)
    int64_t ntiles;
    const void* col[RF_MAXCOLS];
    const uint16_t* valid[RF_MAXCOLS];
    struct RfProbeDev probe[RF_MAXPROBES];
    void* partials;       // Acc per wave-tile (deterministic) or per wave
    void* result;         // final Acc
    uint32_t* counters;   // [1] finished blocks (reset by the last block)
    // Selection: staging (per-wave regions of stage_rows rows), per-wave
    // count and output offset, final columns and count.
    int64_t stage_rows;
    int64_t nwaves;
    int64_t* wave_cnt;
    int64_t* wave_dst;
    void* stage[RF_MAXOUT];
    void* out[RF_MAXOUT];
    int64_t* out_count;
    int64_t out_cap;      // rows the out columns hold (placement stops there)
    // Group join pre-aggregation in LDS: slots per field (groups), and where
    // each block writes its slots ([block][field slot][group]).
    int64_t agg_gpp;
    unsigned long long* agg_part;
    // A group join's probe scan that also counts its joined rows per
    // partition (rank >> part_shift) for the partitioned fold: block b
    // writes its count of partition p at part_boff[p * gridDim.x + b], for
    // the partitions below ceil(groups / 2^part_shift) (groups: part_groups_dev
    // when set, else part_nparts is exact). Null: no counting.
    uint32_t* part_boff;
    int64_t part_nparts;
    const int64_t* part_groups_dev;
    int32_t part_shift;
    int32_t part_pad;
    union RfParam p[RF_MAXPARAMS];
};

// One group join's per-group fold (generated rf_gj_fold): joined rows
// sorted by group, [seg_lo[g], seg_hi[g]) those of group g.
struct RfFoldArgs {
    int64_t groups;
    const int64_t* keys;     // group key (sorted)
    const int64_t* brow;     // build row of the group
    const int64_t* seg_lo;
    const int64_t* seg_hi;
    const void* pay[RF_MAXOUT];   // sorted payload columns (after the rank)
    const void* col[RF_MAXCOLS];  // build table columns read by the outputs
    const uint16_t* valid[RF_MAXCOLS];
    void* out[RF_MAXOUT];         // output columns, one row per group
    union RfParam p[RF_MAXPARAMS];
};

// Device code reads a batch's unit arguments through the constant address
// space (no lane writes them during a launch), so uniform reads become
// scalar loads that later global stores do not force to repeat. The host
// sees a plain pointer of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define RF_UNIT_AS __attribute__((address_space(4)))
#else
#define RF_UNIT_AS
#endif

// Folds of several group join runs in one launch (rf_gj_fold_b): run u
// owns groups [gstart[u], gstart[u+1]) of the launch, its own arguments in
// list[u] (group indexes in list[u] are local to the run).
struct RfFoldBatch {
    const RF_UNIT_AS struct RfFoldArgs* list;
    const int64_t* gstart;  // n + 1 cumulative group counts
    int64_t n;
};

// Concatenation of several runs' joined rows (rank + payload columns) into
// one array; run u's ranks are offset by its first group so one stable sort
// orders every run's groups at once.
struct RfCatSeg {
    const uint32_t* rank;
    const void* pay[8];
    int64_t n;        // joined rows of the run
    int64_t dst;      // first row in the concatenation
    uint32_t base;    // first group of the run
    uint32_t pad;
};

struct RfCatArgs {
    uint32_t* rank;   // concatenated ranks
    void* pay[8];     // concatenated payload columns
    int32_t width[8]; // payload widths in bytes (1, 2, 4, 8)
    int32_t np;
};

// Partitioned aggregation of a group join's joined rows (the default mode):
// the probe scan's staged rows (per-wave regions, rank then payload
// columns) are counted per partition (rank >> shift) and scattered into
// contiguous partitions; rf_gj_agg then folds each partition's groups in LDS.
struct RfPartArgs {
    const uint32_t* rank;      // staged ranks: wave w's rows at [w*stage_rows, +wave_cnt[w])
    const void* pay[8];        // staged payload columns, same layout
    int32_t width[8];          // payload widths in bytes (1, 2, 4, 8)
    int32_t np;
    int32_t shift;             // partition of a row: rank >> shift
    int64_t nparts;
    int64_t stage_rows;
    int64_t nwaves;
    int64_t waves_per_block;   // wave regions one block handles
    const int64_t* wave_cnt;
    int64_t nblocks;           // blocks of the partition kernels
    int64_t chunk;             // rows rf_part_scatter stages in LDS at a time (4096 or 8192)
    // [nparts * nblocks]: rows of partition p in block b at p*nblocks + b,
    // then scanned in place along each partition (b's offset in p).
    uint32_t* boff;
    uint32_t* ptot;            // [nparts] rows per partition
    uint32_t* pstart;          // [nparts + 1] first row of each partition
    uint32_t* prank;           // partitioned ranks
    void* ppay[8];             // partitioned payload columns
    // When set, the number of groups, a count an earlier kernel wrote:
    // partitions from ceil(groups / 2^shift) on are empty (nparts, a bound,
    // still lays out boff, ptot, pstart and the LDS).
    const int64_t* groups_dev;
    int32_t counted;           // the probe scan already wrote boff's counts
    int32_t pad;
};

// The partition of one rf_gj_agg block p: joined rows [pstart[p * stride],
// pstart[(p + 1) * stride]) of prank / RfFoldArgs::pay, groups
// [p << shift, (p + 1) << shift).
struct RfAggArgs {
    const uint32_t* prank;
    const uint32_t* pstart;
    int64_t stride;
    int32_t shift;
    int32_t pad;
    // When set, the number of groups (RfFoldArgs::groups is then a bound):
    // a count an earlier kernel wrote.
    const int64_t* groups_dev;
    // rf_gj_lds_done: the probe scan's block slots ([nblk][field slot][gpp]).
    const unsigned long long* bpart;
    int64_t nblk;
    int64_t gpp;
};

// Batched execution: a persistent launch over (work unit, wave-tile) items.
struct RfBatchHdr {
    int64_t nunits;      // work units of this launch
    int64_t nitems;      // their total wave-tiles
    int64_t unit_base;   // first unit in the round's work array
    int64_t start_base;  // first entry in the round's item_start array
};

// Pointers are fixed per runner (a captured launch replays unchanged); the
// header, written per round, says where this launch's units are.
struct RfBatchArgs {
    const struct RfBatchHdr* hdr;
    const RF_UNIT_AS struct RfScanArgs* work;  // the round's units
    const int64_t* item_start;         // per launch: nunits + 1 cumulative tile counts
};

// One run reduction (generated rf_seg_fold): run r covers rows
// [run_ends[r-1], run_ends[r]) (from 0 for r = 0), clamped to `rows`.
struct RfSegArgs {
    int64_t runs;
    int64_t rows;
    int64_t key0;                 // ranged folds: the key of fold 0
    const int32_t* run_ends;
    const void* keys;             // run values
    const void* col[RF_MAXCOLS];  // columns the row code and outputs read
    const uint16_t* valid[RF_MAXCOLS];
    void* out[RF_MAXOUT];         // output columns, one row per run
    struct RfProbeDev probe[RF_MAXPROBES];  // tables read at the run's key
    union RfParam p[RF_MAXPARAMS];
};
