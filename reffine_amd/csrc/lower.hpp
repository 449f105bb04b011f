// GPU lowering of a reffine Func onto the scan kernel skeleton.
//
// The reference lowers Reduce/Op nodes to a sequential Loop
// (src/pass/loopgen.cpp:44-269) that folds one row at a time into a struct
// state. The GPU counterpart keeps the same row semantics but re-associates
// the fold: every flattened state field must update as
//     s_f' = pred_f(row) ? s_f (+|*|min|max|&&|||) contrib_f(row) : s_f
// with pred_f and contrib_f free of state, so rows can be combined in a
// tree. Such fields are produced by the reference's own idioms: _sel(pred,
// new_state, s) (benchmark/include/tpch.h:174-208), per-field _sel
// (tpch.h:412-478) and plain accumulation (test/src/test_reduce.cpp:61-122).
// A field that does not decompose makes the query RF_ERR_UNSUPPORTED.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ir.hpp"

namespace rf {

// Row-level leaf symbols produced by the lowering (Sym nodes whose names
// start with '$'):
//   $col.T.C    value of column C of func input T at the current row
//               (T is the scanned table) or at the probed row (T probed)
//   $valid.T.C  validity bit of that column
//   $found.T    the probe into table T found the key
//   $row        the row index (or the iterator value of a range scan)
//   $arg.I      scalar func input I
//   $len.T.C    length of column C of table T (host-known)
//   $s.F        state field F (only inside the decomposition)
std::string leaf_name(const N& n);  // "" if not a leaf
bool is_leaf(const N& n, const char* prefix);

enum class AggOp { ADD, MUL, MIN, MAX, AND, OR };

struct Field {
    Type type;           // primitive type of the flattened state field
    AggOp op = AggOp::ADD;
    bool updated = false;
    N pred;              // bool row predicate (never null when updated)
    N contrib;           // contribution (field type)
    N init;              // init expression of the field (host-evaluated)
    std::vector<int> path;  // position inside the nested state struct
};

struct ReduceComp {
    Type state_type;
    std::vector<Field> fields;
    std::string sym_name;  // table symbol holding this reduction
    // A fold the kernels cannot re-associate (a field's update is not
    // `pred ? s (+) x : s`, or reads another field): the state is folded row
    // by row in order, s = pred(row) ? next(s, row) : s, every field at once
    // (`$s.F` leaves are the current state).
    bool sequential = false;
    N seq_pred;                 // row predicate (null: every row)
    std::vector<N> seq_next;    // the next value of each flattened field
};

struct OpComp {
    std::vector<N> values;   // iterator values then outputs (row leaves)
    std::vector<Type> types;
    int n_iters = 1;
    N pred;                  // row qualifies (bool)
    Type vec_type;           // reffine type of the output vector
    std::string sym_name;
};

struct Probe {
    int table;  // func input index of the probed table
    N key;      // int64 row expression giving the key
    N guard;    // probe only rows where this holds (null: every row)
    bool external = false;  // uses an index supplied at run time (group join)
};

struct GroupJoinPlan;
struct SegReducePlan;
struct RunStagePlan;

struct Plan {
    enum Source { TABLE_ROWS, RANGE, MERGE } source = TABLE_ROWS;
    // MERGE: an Op over _in(t, A) | _in(t, B) walks the merge of the two
    // tables' key columns (row leaves $mkey, $mcol.S.c, $mvalid.S.c with
    // S = 0 for A, 1 for B). A side given as _in(t + c, X) (c constant) is
    // merged by key + c (merge_off[S]; wrapping int64 addition).
    int merge_a = -1, merge_b = -1;
    int64_t merge_off[2] = {0, 0};
    std::string func_name;
    std::vector<N> inputs;     // func inputs (syms)
    int driver = -1;           // func input index of the scanned table
    N start, end;              // host expressions: rows [start, end)
    N row_pred;                // common row predicate (validity, bounds)
    std::vector<Probe> probes;
    std::vector<ReduceComp> reduces;
    std::vector<OpComp> ops;   // at most one
    N output;                  // func output expression over component syms
    Type output_type;
    std::vector<std::pair<N, N>> tbl;  // the func's symbol table
    bool vectorize = false;
    // Selection outputs are evaluated in the scan from the loaded rows and
    // staged (early materialisation) instead of read again by row index in
    // the placement kernel; for selections that keep a large share of rows.
    bool early_sel = false;
    // External probes read a rank map (a group join's dense build keys):
    // the kernel calls the branch-free rank lookup; it must only be given
    // rank maps.
    bool dense_probes = false;
    // The function is one sequential fold (ReduceComp::sequential): a
    // single wave folds the rows in order (rf_qs).
    bool sequential = false;
    // Independent components run as queries of their own
    // (split_components); this plan holds only the inputs, the symbol
    // table and the output expression over the components.
    bool composite = false;
    std::shared_ptr<GroupJoinPlan> gj;  // set: the function is a group join
    // Set on a group join's probe plan compiled for pre-aggregation in LDS:
    // the scan folds each joined row into its block's LDS slots (one per
    // group and updated field of *lds_agg) instead of staging it.
    const GroupJoinPlan* lds_agg = nullptr;
    std::shared_ptr<SegReducePlan> seg; // set: a reduction per run of a run-end column
    std::shared_ptr<RunStagePlan> runs; // set: a runs table is built first (T's place)

    std::string describe() const;
};

// A hash join with a grouped aggregate, written in reffine as an Op over the
// build table B whose outputs reduce the rows of the probe table P whose key
// equals the Op's iterator:
//   _op([t], _in(t, B) & bpred(B[t]),
//       [f(t, B[t], _red(_op([i], _in(i, P) & (key(P[i]) == t) & ppred(P[i]), outs), init, acc))])
// The reference runs the inner reduction once per build row
// (src/pass/loopgen.cpp:113-269). The GPU plan: select the qualifying build
// keys and rows, sort them by key (ranks), index key -> rank, scan P once
// selecting (rank, per-field contributions) of joined rows, stable-sort those
// by rank and fold each group in ascending probe-row order from the
// reduction's init -- the inner loop's fold order.
struct GField {
    Type type;
    AggOp op = AggOp::ADD;
    bool updated = false;
    N init;            // host expression
    N const_contrib;   // constant contribution (no payload column)
    int pay = -1;      // payload column of the contribution
    int pred_pay = -1; // payload column of the field's own predicate
    std::vector<int> path;
};

struct GroupJoinPlan {
    int build = -1, probe = -1;  // func input indexes of B and P
    Plan build_plan;             // Op over B: (key i64, build row i64)
    Plan probe_plan;             // Op over P: (rank i32, payload...); probe 0 is the group index
    std::vector<GField> fields;  // flattened state of the inner reduction
    Type state_type;
    std::vector<Type> pay_types; // payload columns after the rank
    std::vector<N> out_values;   // $key, then the outer Op outputs over $key, $acc.F, B's row
    std::vector<Type> out_types;
    Type vec_type;
    std::string sym_name;
};

// A computed vector the function reads as a table (an Op, or an Op with an
// index built over it, named in the symbol table and used through Element,
// In, SubVector, ReadData, Locate or Length): run first as a function of its
// own (same inputs), its result passed to the function as one more input.
struct Prelude {
    N func;            // the sub-function (output: the Op)
    bool index = false; // build_index on its result (BuildIndex)
};

// Splits the computed vectors out of `func`: returns the function with one
// more input per prelude (appended in order) and fills `pre`; returns
// `func` itself when it reads none.
N split_materialized(const N& func, std::vector<Prelude>& pre);

// Lowers `func` (a parsed IR Func) into a plan; throws Unsupported for
// shapes the kernels do not cover and IRError for ill-formed IR.
Plan lower_func(const N& func, bool vectorize);

// The components (Reduce / Op symbols, through the symbol table) the
// function's output combines, in the order the output names them.
std::vector<N> output_components(const N& func);

// A function whose components cannot share one scan (they read different
// tables, or one is a sequential fold): one function per component, each
// with that component as its output, on the same inputs and symbol table.
// The reference emits each component as a loop of its own
// (src/pass/loopgen.cpp:113-269), one after another.
std::vector<N> split_components(const N& func);

// Evaluation of host-side expressions (bounds, init values, the func output
// over reduction results).
struct HostVal {
    Type type;
    bool is_float = false;
    double f = 0;
    int64_t i = 0;
    std::vector<HostVal> fields;  // struct
    void* table = nullptr;        // vector-typed output (device table)
};

}  // namespace rf

namespace rf {

// An Op over the runs of the run-end encoded column 0 of a 2-dimensional
// table whose outputs reduce each run's rows:
//   _op([t], _in(t, T), [f(t, _red(T[t], init, acc))])
// (test/src/test_multidim.cpp:9-30). T[t] is the sub-vector of the run's
// rows (src/pass/loopgen.cpp:34-38); the reduction folds them in row order
// with the SubVector body condition, the validity of column 1
// (loopgen.cpp:221-242). One output row per run, in run order.
struct SegReducePlan {
    int table = -1;             // func input index of T
    Plan sub;                   // the row code's plan: driver T, probes keyed by $key
    // Ranged form: one fold per value $key of a range iterator lb <= $key <
    // ub (host expressions), over rows [lo, hi) of T (expressions of $key)
    // instead of T's runs.
    bool ranged = false;
    N lb, ub, lo, hi;
    std::vector<Field> fields;  // row predicates/contributions over $col.T.C, $valid.T.C
    Type state_type;
    Type key_type;              // type of the run values (the iterator)
    std::vector<N> out_values;  // $key, then the Op outputs over $key and $acc.F
    std::vector<Type> out_types;
    Type vec_type;
    std::string sym_name;
};

// The first stage of a function with reductions over the runs of a
// run-end table T used by an Op over In(t, T): folds every run into a runs
// table (key, then every reduction's fields) that the function's plan
// scans in T's place.
struct RunStagePlan {
    int table = -1;       // func input index of T
    N input;              // T's symbol with its own (run-end) type
    SegReducePlan seg;    // out_values: $key, then $acc.F for every field
};

}  // namespace rf
