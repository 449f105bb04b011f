// Emits the per-query row code (`struct Q`) for the scan skeleton in
// kernels/rf_device.h. The emitted expressions keep the C text semantics of
// the reference's CEmitter (src/pass/cemitter.cpp:80-170): operators are
// printed infix on values of the IR's C types, so integer promotion, casts
// and comparisons behave as in the CPU JIT; kernels are compiled with
// -ffp-contract=off so that no multiply-add is fused.
#pragma once

#include <string>
#include <vector>

#include "lower.hpp"

namespace rf {

struct ParamSlot {
    enum Kind { CONST, ARG, LEN } kind;
    Type type;
    double value = 0;  // CONST
    int arg = 0;       // ARG: func input index; LEN: table input index
    int col = 0;       // LEN: column
};

struct ColSlot {
    int table;       // func input index
    int col;         // column in that table
    Type type;
    bool nullable;   // validity bitmap present (specialisation)
    bool vector;     // loaded with the scan's vector loads (driver column)
};

struct AccField {
    int comp;        // reduce component
    int field;       // field index within the component
    Type type;
    AggOp op;
};

struct Tuning {
    int block = 256;
    int items = 4;
    int sub = 4;
    int waves_per_cu = 16;  // launch width: resident waves aimed at per CU
    int pipeline = 0;       // 1: overlap the next load group's loads with this one's rows
    int lsub = 0;           // sub-steps per load group (divides sub); 0: from the load budget
    // A group join's hash index over its build keys (sparse keys): capacity
    // >= index_scale * keys, rounded up to a power of two (load <= 1/scale).
    int index_scale = 2;
};

struct KernelSource {
    std::string code;              // Q struct + kernel entry
    std::string entry = "rf_q";
    std::vector<ParamSlot> params;
    std::vector<ColSlot> cols;
    std::vector<AccField> acc;     // Acc field order
    std::vector<int> acc_offset;   // byte offset of each Acc field
    int acc_bytes = 0;
    bool has_sel = false;
    int lds_agg_slots = 0;         // a probe that pre-aggregates in LDS: slots per group
    bool lds_emit = false;         // selection compacted in LDS (can count partitions)
    std::vector<Type> out_types;   // selection output columns
    std::vector<Type> stage_types; // staged per selected row: row index, carried probe results
    Tuning tune;
    bool det = false;
    // A sequential fold (rf_qs, one wave, grid 1): the Acc fields are the
    // final state, not partial folds to combine with the init.
    bool sequential = false;
    // Run folds: rows per block the LDS copy holds (0: not staged), and the
    // capacity of the rf_seg_fold_long variant (0: none).
    int cap = 0;
    int alt_cap = 0;
};

// `nullable[t][c]` says whether column c of func input t has a bitmap.
// `pack_probes[j]` as for gen_seg_fold: probe j's table is a packed rank map
// over ascending keys with at least one key (probed without branches).
KernelSource gen_scan(const Plan& plan,
                      const std::vector<std::vector<bool>>& nullable,
                      const Tuning& tune, bool det,
                      const std::vector<bool>& pack_probes = {});

// Per-group fold and output kernel (rf_gj_fold) of a group join;
// `nullable` describes the function inputs as for gen_scan.
// Whether a group join's fields fold with LDS atomics in the default mode
// (rf_gj_agg): sums of integers or floats, integer min / max.
bool gj_agg_ok(const GroupJoinPlan& gj);
std::string gj_slot_identity(const GField& gf);
int gj_lds_done_waves(int nslot);
KernelSource gen_group_fold(const GroupJoinPlan& gj,
                            const std::vector<std::vector<bool>>& nullable);

// Per-run fold and output kernel (rf_seg_fold) of a run reduction.
// `pack_probes[j]`: probe j's table is a packed rank map over ascending keys
// with at least one key (the row loop then probes it without branches).
KernelSource gen_seg_fold(const SegReducePlan& sp, const Plan& plan,
                          const std::vector<std::vector<bool>>& nullable,
                          const std::vector<bool>& pack_probes = {});

// Picks ITEMS from the narrowest driver column the row code loads.
Tuning default_tuning(const Plan& plan);
// The same, for the deterministic mode when `det` (larger wave-tiles).
Tuning default_tuning(const Plan& plan, bool det);
// Sub-steps one set of loads may hold in registers.
int load_cap(const Plan& plan, int items, bool pipelined);

// Identity of a fold as a C literal of `t`.
std::string fold_identity(AggOp op, const Type& t);

}  // namespace rf
