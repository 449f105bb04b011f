// Row-code emitter for the scan skeleton (see hipgen.hpp).
#include "hipgen.hpp"

#include "kernels/rf_abi.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <functional>
#include <sstream>

namespace rf {

namespace {

std::string cty(const Type& t) { return t.cstr(); }

std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == d) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    out.push_back(cur);
    return out;
}

// ROWS: the scan's row code over vector-loaded rows; MOVE: the outputs of
// one qualifying row, read by index where the placement kernel moves the
// row (probe results the row code computed come from staging); FINAL: one
// group of a group join (build-table columns at the group's build row).
// ROWS: a scan's row block (columns in registers per item); MOVE: one row
// read by index (placement, run folds); FINAL: a group's outputs; SEQ: the
// row of a sequential fold, its columns broadcast from the lane that loaded
// them (c<slot> / v<slot>), the state in s<field>.
// MERGE: the current rows of a union merge (side 0 at row i, side 1 at row
// j, ok0 / ok1 whether each is inside its table; key the iterator).
enum class Mode { ROWS, MOVE, FINAL, SEQ, MERGE };

struct Gen {
    const Plan& plan;
    const std::vector<std::vector<bool>>& nullable;
    KernelSource& ks;
    std::map<std::string, int> param_ix;
    std::map<std::pair<int, int>, int> col_ix;
    std::set<int> carried;  // probes the row code computes; emit reuses them
    // Run folds: scanned-table columns read from the block's LDS copy
    // (sm<slot>[row - blo]) instead of global memory; the slots so read.
    bool lds_rows = false;
    std::set<int> lds_slots;
    // Run folds: probes whose table is a packed rank map with ascending keys
    // and at least one key; their row code probes (rf_probe_pack) and reads
    // the found row (clamped to row 0, then selected) without branches.
    std::vector<bool> pack_probes;
    // Sequential folds: the column slots (values, validity) the row reads.
    std::set<int> seq_cols, seq_valid;
    bool packed(int j) const {
        return j < (int)pack_probes.size() && pack_probes[j];
    }

    Gen(const Plan& p, const std::vector<std::vector<bool>>& nl, KernelSource& k)
        : plan(p), nullable(nl), ks(k) {}

    bool is_nullable(int t, int c) const {
        return t < (int)nullable.size() && c < (int)nullable[t].size() && nullable[t][c];
    }

    int col_slot(int t, int c, bool vec) {
        auto key = std::make_pair(t, c);
        auto it = col_ix.find(key);
        if (it != col_ix.end()) {
            if (vec) ks.cols[it->second].vector = true;
            return it->second;
        }
        if ((int)ks.cols.size() >= 32) throw Unsupported("query reads more than 32 columns");
        ColSlot s;
        s.table = t;
        s.col = c;
        s.type = plan.inputs[t]->type.dtypes[c];
        s.nullable = is_nullable(t, c);
        s.vector = vec;
        ks.cols.push_back(s);
        int ix = (int)ks.cols.size() - 1;
        col_ix[key] = ix;
        return ix;
    }

    std::string param(const ParamSlot& p, const std::string& key) {
        int ix;
        auto it = param_ix.find(key);
        if (it != param_ix.end()) {
            ix = it->second;
        } else {
            if ((int)ks.params.size() >= 96) throw Unsupported("more than 96 constants");
            ks.params.push_back(p);
            ix = (int)ks.params.size() - 1;
            param_ix[key] = ix;
        }
        const Type& t = p.type;
        std::string s = "L.p[" + std::to_string(ix) + "]";
        if (t.bt == BT::BOOL) return "(" + s + ".i != 0)";
        if (t.is_float()) return "((" + cty(t) + ")" + s + ".f)";
        return "((" + cty(t) + ")" + s + ".i)";
    }

    std::string const_param(const N& c) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "c:%s:%a", c->type.str().c_str(), c->val);
        ParamSlot p{ParamSlot::CONST, c->type, c->val, 0, 0};
        return param(p, buf);
    }
};

// Expression emitter with structural CSE into `const auto` temporaries.
struct Emitter {
    Gen& g;
    Mode mode;
    std::ostringstream& os;
    std::string indent;
    std::map<std::string, std::string> memo;
    std::set<int> probes_done;
    int ntmp = 0;

    Emitter(Gen& gen, Mode m, std::ostringstream& o, std::string ind)
        : g(gen), mode(m), os(o), indent(std::move(ind)) {}

    std::string tmp(const std::string& rhs) {
        std::string v = "t" + std::to_string(ntmp++);
        os << indent << "const auto " << v << " = " << rhs << ";\n";
        return v;
    }

    std::set<int> staged;  // MOVE: carried probe results read from staging

    std::string probe(int j) {
        std::string h = "h" + std::to_string(j);
        if (mode == Mode::MOVE && g.carried.count(j)) {
            staged.insert(j);
            return "sc" + std::to_string(j);
        }
        if (!probes_done.count(j)) {
            const Probe& pr = g.plan.probes[j];
            std::string k = expr(pr.key);
            if (g.packed(j)) {
                const std::string gd = pr.guard ? expr(pr.guard) : std::string("true");
                os << indent << "const int64_t " << h << " = rf_probe_pack(L.probe[" << j
                   << "], (int64_t)" << k << ", " << gd << ");\n";
            } else if (pr.external && g.plan.dense_probes) {
                const std::string gd = pr.guard ? expr(pr.guard) : std::string("true");
                os << indent << "const int64_t " << h << " = rf_probe_dense(L.probe[" << j
                   << "], (int64_t)" << k << ", " << gd << ");\n";
            } else if (pr.guard) {
                std::string gd = expr(pr.guard);
                os << indent << "const int64_t " << h << " = (" << gd << ") ? rf_probe(L.probe["
                   << j << "], (int64_t)" << k << ") : (int64_t)-1;\n";
            } else {
                os << indent << "const int64_t " << h << " = rf_probe(L.probe[" << j
                   << "], (int64_t)" << k << ");\n";
            }
            if (mode == Mode::ROWS) os << indent << "cy.p" << j << "[s][k] = " << h << ";\n";
            probes_done.insert(j);
        }
        return h;
    }

    std::string leaf_text(const N& e) {
        const std::string name = leaf_name(e);
        auto parts = split(name, '.');
        const std::string& kind = parts[0];
        if (kind == "$mcol" || kind == "$mvalid") {
            if (mode != Mode::MERGE) throw IRError("merge leaf outside a union merge");
            const int S = std::stoi(parts[1]), c = std::stoi(parts[2]);
            const int t = S == 0 ? g.plan.merge_a : g.plan.merge_b;
            const std::string I = S == 0 ? "i" : "j", OK = S == 0 ? "ok0" : "ok1";
            if (kind == "$mvalid" && !g.is_nullable(t, c)) return OK;
            const int sl = g.col_slot(t, c, false);
            const std::string SL = std::to_string(sl);
            if (kind == "$mvalid")
                return tmp("(" + OK + " && rf_valid_bit(L.valid[" + SL + "], " + I + "))");
            // A side past its last row reads 0.
            const std::string T = cty(g.ks.cols[sl].type);
            return tmp("(" + OK + " ? rf_g((const " + T + "*)L.col[" + SL + "])[" + I + "] : (" +
                       T + ")0)");
        }
        if (kind == "$mkey") {
            if (mode != Mode::MERGE) throw IRError("merge key outside a union merge");
            return "key";
        }
        if (kind == "$col") {
            int t = std::stoi(parts[1]), c = std::stoi(parts[2]);
            if (t != g.plan.driver) throw IRError("column leaf of a non-scanned table");
            int s = g.col_slot(t, c, mode == Mode::ROWS);
            const std::string T = cty(g.ks.cols[s].type);
            if (mode == Mode::ROWS) return "c" + std::to_string(s) + ".v[k]";
            if (mode == Mode::SEQ) {
                g.seq_cols.insert(s);
                return "c" + std::to_string(s);
            }
            if (mode == Mode::FINAL) return "rf_g((const " + T + "*)L.col[" + std::to_string(s) + "])[brow]";
            if (g.lds_rows) {
                g.lds_slots.insert(s);
                return "sm" + std::to_string(s) + "[row - blo]";
            }
            return "rf_g((const " + T + "*)L.col[" + std::to_string(s) + "])[row]";  // MOVE
        }
        if (kind == "$valid") {
            int t = std::stoi(parts[1]), c = std::stoi(parts[2]);
            if (t != g.plan.driver) throw IRError("validity leaf of a non-scanned table");
            if (!g.is_nullable(t, c)) return "true";
            int s = g.col_slot(t, c, mode == Mode::ROWS);
            if (mode == Mode::ROWS) return "(((vb" + std::to_string(s) + " >> k) & 1u) != 0u)";
            if (mode == Mode::SEQ) {
                g.seq_valid.insert(s);
                return "v" + std::to_string(s);
            }
            if (mode == Mode::FINAL) return "rf_valid_bit(L.valid[" + std::to_string(s) + "], brow)";
            return "rf_valid_bit(L.valid[" + std::to_string(s) + "], row)";
        }
        if (kind == "$pcol") {
            int j = std::stoi(parts[1]), c = std::stoi(parts[2]);
            int t = g.plan.probes[j].table;
            int s = g.col_slot(t, c, false);
            std::string h = probe(j);
            const std::string T = cty(g.ks.cols[s].type);
            if (g.packed(j)) {
                const std::string v = tmp("rf_g((const " + T + "*)L.col[" + std::to_string(s) +
                                          "])[" + h + " < 0 ? 0 : " + h + "]");
                return tmp("(" + h + " >= 0 ? " + v + " : (" + T + ")0)");
            }
            return tmp("(" + h + " >= 0 ? rf_g((const " + T + "*)L.col[" + std::to_string(s) +
                       "])[" + h + "] : (" + T + ")0)");
        }
        if (kind == "$found") {
            int j = std::stoi(parts[1]);
            return "(" + probe(j) + " >= 0)";
        }
        if (kind == "$prow") return probe(std::stoi(parts[1]));
        if (kind == "$row") {
            if (mode == Mode::FINAL) throw IRError("row index in a group join output");
            if (mode == Mode::MERGE) throw IRError("row index in a union merge");
            return "row";
        }
        if (kind == "$key") return "key";
        if (kind == "$acc") return "s" + parts[1];
        if (kind == "$s" && mode == Mode::SEQ) return "s" + parts[1];
        if (kind == "$arg") {
            ParamSlot p{ParamSlot::ARG, e->type, 0, std::stoi(parts[1]), 0};
            return g.param(p, name + ":" + e->type.str());
        }
        if (kind == "$len") {
            ParamSlot p{ParamSlot::LEN, e->type, 0, std::stoi(parts[1]), std::stoi(parts[2])};
            return g.param(p, name);
        }
        throw IRError("unexpected leaf " + name + " in row code");
    }

    std::string expr(const N& e) {
        if (e->kind == K::Sym) {
            auto it = memo.find(e->name);
            if (it != memo.end()) return it->second;
            std::string r = leaf_text(e);
            memo[e->name] = r;
            return r;
        }
        if (e->kind == K::Const) return g.const_param(e);
        std::string key = to_text(e);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        std::string r;
        switch (e->kind) {
            case K::Cast:
                r = tmp("(" + cty(e->type) + ") " + expr(e->args[0]));
                break;
            case K::Select:
                r = tmp("(" + expr(e->args[0]) + " ? " + expr(e->args[1]) + " : " +
                        expr(e->args[2]) + ")");
                break;
            case K::Nary:
                r = tmp(nary(e));
                break;
            case K::ReadData: {
                // Column c of table t at a computed row; 0 outside the table.
                const std::string tn = leaf_name(e->args[0]);
                if (tn.rfind("$tab.", 0) != 0) throw IRError("ReadData of " + tn);
                const int t = std::stoi(tn.substr(5)), c = (int)e->idx;
                const int s = g.col_slot(t, c, false);
                const std::string T = cty(g.ks.cols[s].type);
                const std::string i = "(int64_t)" + expr(e->args[1]);
                ParamSlot p{ParamSlot::LEN, types::IDX_T, 0, t, c};
                const std::string len =
                    g.param(p, "$len." + std::to_string(t) + "." + std::to_string(c));
                r = tmp("(" + i + " >= 0 && " + i + " < " + len + ") ? rf_g((const " + T +
                        "*)L.col[" + std::to_string(s) + "])[" + i + "] : (" + T + ")0");
                break;
            }
            default:
                throw Unsupported("node in row code: " + key);
        }
        memo[key] = r;
        return r;
    }

    std::string nary(const N& e) {
        auto a = [&](int i) { return expr(e->args[i]); };
        const std::string T = cty(e->type);
        switch (e->op) {
            case MathOp::ADD: return "(" + a(0) + " + " + a(1) + ")";
            case MathOp::SUB: return "(" + a(0) + " - " + a(1) + ")";
            case MathOp::MUL: return "(" + a(0) + " * " + a(1) + ")";
            case MathOp::DIV: return "(" + a(0) + " / " + a(1) + ")";
            case MathOp::MOD: return "(" + a(0) + " % " + a(1) + ")";
            case MathOp::MAX:
                return "rf_max<" + T + ">((" + T + ")" + a(0) + ", (" + T + ")" + a(1) + ")";
            case MathOp::MIN:
                return "rf_min<" + T + ">((" + T + ")" + a(0) + ", (" + T + ")" + a(1) + ")";
            case MathOp::ABS: return "rf_abs<" + T + ">((" + T + ")" + a(0) + ")";
            case MathOp::NEG: return "(-" + a(0) + ")";
            case MathOp::SQRT: return "sqrt(" + a(0) + ")";
            case MathOp::POW: return "pow(" + a(0) + ", " + a(1) + ")";
            case MathOp::CEIL: return "ceil(" + a(0) + ")";
            case MathOp::FLOOR: return "floor(" + a(0) + ")";
            case MathOp::LT: return "(" + a(0) + " < " + a(1) + ")";
            case MathOp::LTE: return "(" + a(0) + " <= " + a(1) + ")";
            case MathOp::GT: return "(" + a(0) + " > " + a(1) + ")";
            case MathOp::GTE: return "(" + a(0) + " >= " + a(1) + ")";
            case MathOp::EQ: return "(" + a(0) + " == " + a(1) + ")";
            case MathOp::NOT: return "(!" + a(0) + ")";
            case MathOp::AND: return "(" + a(0) + " && " + a(1) + ")";
            case MathOp::OR: return "(" + a(0) + " || " + a(1) + ")";
        }
        throw IRError("bad operator");
    }
};

std::string combine_expr(AggOp op, const Type& t, const std::string& a,
                         const std::string& b) {
    const std::string T = cty(t);
    switch (op) {
        case AggOp::ADD: return "(" + a + " + " + b + ")";
        case AggOp::MUL: return "(" + a + " * " + b + ")";
        case AggOp::MIN: return "rf_min<" + T + ">(" + a + ", (" + T + ")" + b + ")";
        case AggOp::MAX: return "rf_max<" + T + ">(" + a + ", (" + T + ")" + b + ")";
        case AggOp::AND: return "(" + a + " && " + b + ")";
        case AggOp::OR: return "(" + a + " || " + b + ")";
    }
    return "?";
}

}  // namespace

std::string fold_identity(AggOp op, const Type& t) {
    const std::string T = t.cstr();
    switch (op) {
        // -0.0 is the exact additive identity of IEEE floats (+0.0 + -0.0 is
        // +0.0, so a fold from +0.0 would lose a sum's -0.0).
        case AggOp::ADD: return t.is_float() ? "(" + T + ")(-0.0)" : "(" + T + ")0";
        case AggOp::MUL: return "(" + T + ")1";
        case AggOp::AND: return "true";
        case AggOp::OR: return "false";
        case AggOp::MIN:
        case AggOp::MAX: {
            bool mx = op == AggOp::MAX;  // identity of max is the lowest value
            if (t.is_float()) return mx ? "(" + T + ")(-__builtin_inf())" : "(" + T + ")__builtin_inf()";
            switch (t.bt) {
                case BT::BOOL: return mx ? "false" : "true";
                case BT::INT8: return mx ? "(int8_t)(-128)" : "(int8_t)127";
                case BT::INT16: return mx ? "(int16_t)(-32768)" : "(int16_t)32767";
                case BT::INT32: return mx ? "(int32_t)(-2147483647 - 1)" : "(int32_t)2147483647";
                case BT::INT64: case BT::IDX:
                    return mx ? "(int64_t)(-9223372036854775807LL - 1)" : "(int64_t)9223372036854775807LL";
                case BT::UINT8: return mx ? "(uint8_t)0" : "(uint8_t)255";
                case BT::UINT16: return mx ? "(uint16_t)0" : "(uint16_t)65535";
                case BT::UINT32: return mx ? "(uint32_t)0" : "(uint32_t)4294967295u";
                case BT::UINT64: return mx ? "(uint64_t)0" : "(uint64_t)18446744073709551615ull";
                default: break;
            }
        }
    }
    throw IRError("no identity for fold");
}

static void collect_driver_cols(const N& e, const Plan& plan, std::map<int, int>& cols) {
    if (!e) return;
    if (is_leaf(e, "$col.")) {
        auto parts = split(leaf_name(e), '.');
        if (std::stoi(parts[1]) == plan.driver) cols[std::stoi(parts[2])] = e->type.width();
    }
    for (auto& a : e->args) collect_driver_cols(a, plan, cols);
}

static void driver_row_shape(const Plan& plan, int& wmax, int& row_bytes, int& nacc) {
    std::map<int, int> cols;  // driver column -> width, loaded per row
    for (auto& r : plan.reduces)
        for (auto& f : r.fields)
            if (f.updated) {
                collect_driver_cols(f.pred, plan, cols);
                collect_driver_cols(f.contrib, plan, cols);
            }
    for (auto& o : plan.ops) collect_driver_cols(o.pred, plan, cols);
    for (auto& p : plan.probes) {
        collect_driver_cols(p.key, plan, cols);
        collect_driver_cols(p.guard, plan, cols);
    }
    wmax = row_bytes = nacc = 0;
    for (auto& kv : cols) {
        wmax = std::max(wmax, kv.second);
        row_bytes += kv.second;
    }
    for (auto& r : plan.reduces)
        for (auto& f : r.fields) nacc += f.updated ? 1 : 0;
}

// Sub-steps one set of loads may hold: about 640 (pipelined) or 320 bytes
// per lane, at most 32 rows per lane.
int load_cap(const Plan& plan, int items, bool pipelined) {
    int wmax, row_bytes, nacc;
    driver_row_shape(plan, wmax, row_bytes, nacc);
    // As default_tuning's budget: a pipelined state of more than 8
    // accumulators keeps half the load bytes in flight (two register sets
    // of a 36-accumulator state at 640 bytes spilled).
    const int budget = pipelined ? (nacc <= 8 ? 640 : 320) : 320;
    int sub = 4;
    if (row_bytes > 0) sub = std::max(1, std::min(16, budget / (items * row_bytes)));
    sub = std::min(sub, std::max(1, 32 / items));
    if (!plan.probes.empty()) sub = std::max(1, 8 / items);
    return sub;
}

Tuning default_tuning(const Plan& plan, bool det) {
    Tuning t = default_tuning(plan);
    if (det) {
        // The deterministic mode combines each wave-tile across lanes; tiles
        // of several load groups (at least 4096 rows, 8192 with more than 16
        // accumulators: the combine costs per accumulator) amortise that
        // combine (measured on the 60M-row grouped scan, 36 accumulators:
        // 0.417 ms at 8192 rows per tile, 0.460 ms at 4096).
        int nacc = 0;
        for (auto& r : plan.reduces)
            for (auto& f : r.fields) nacc += f.updated ? 1 : 0;
        const int want = nacc > 16 ? 8192 : 4096;
        const int rows = t.sub * 64 * t.items;
        const int groups = std::max(1, (want + rows - 1) / rows);
        t.sub *= groups;
    }
    return t;
}

Tuning default_tuning(const Plan& plan) {
    std::map<int, int> cols;  // driver column -> width, loaded per row
    for (auto& r : plan.reduces)
        for (auto& f : r.fields)
            if (f.updated) {
                collect_driver_cols(f.pred, plan, cols);
                collect_driver_cols(f.contrib, plan, cols);
            }
    for (auto& o : plan.ops) collect_driver_cols(o.pred, plan, cols);
    for (auto& p : plan.probes) {
        collect_driver_cols(p.key, plan, cols);
        collect_driver_cols(p.guard, plan, cols);
    }
    Tuning t;
    int wmax = 0, row_bytes = 0;
    for (auto& kv : cols) {
        wmax = std::max(wmax, kv.second);
        row_bytes += kv.second;
    }
    // ITEMS rows per lane make the widest column one dense 16-byte load per
    // lane per sub-step (narrower columns get dense 8/4/2-byte loads); SUB
    // sub-steps put about 256 bytes per lane in flight per wave-tile. With
    // contiguous wave ranges at 16 waves per CU this is the streaming shape
    // that saturates HBM on gfx950.
    t.items = wmax >= 8 ? 2 : (wmax == 4 ? 4 : (wmax == 2 ? 8 : 16));
    if (wmax == 0 || plan.source == Plan::RANGE) t.items = 4;
    // Wave-tiles are double-buffered (the next tile's loads overlap this
    // tile's rows) at 8 waves per CU; a very large state (more than 48
    // accumulators) keeps one register set at 16 waves per CU.
    int nacc = 0;
    for (auto& r : plan.reduces)
        for (auto& f : r.fields) nacc += f.updated ? 1 : 0;
    // Up to 16 accumulators (the five 3-field buckets of TPC-DS Q9) pipeline
    // with half the load bytes in flight: 0.143 ms against 0.188 ms
    // unpipelined on 30M rows. So does the 36-accumulator grouped scan (six
    // Q1 groups of six fields, 38-byte rows): 349 us against 369 us at 16
    // waves per CU unpipelined on 60M rows (one predicate branch per group
    // and DPP reductions made room for the second register set).
    t.pipeline = nacc <= 48 ? 1 : 0;
    t.waves_per_cu = t.pipeline ? 8 : 16;
    int budget = t.pipeline ? (nacc <= 8 ? 640 : 320) : 320;  // load bytes per lane per tile
    if (t.pipeline && nacc <= 8) {
        // A small state streams fastest with three waves per CU (one block
        // of 192 threads, ~480 load bytes per lane per group, two register
        // sets): on the 60M-row C2 scan 194-196 us against 212 us at four
        // waves (256 threads) and 235 us at two; five and more waves per CU
        // measured slower still (more concurrent streams, 281 us at eight).
        t.block = 192;
        t.waves_per_cu = 3;
        budget = 480;
    }
    t.sub = 4;
    if (row_bytes > 0) t.sub = std::max(1, std::min(16, budget / (t.items * row_bytes)));
    // At most 32 rows per lane per wave-tile: narrow columns would otherwise
    // unpack into more registers than a wave has.
    t.sub = std::min(t.sub, std::max(1, 32 / t.items));
    if (!plan.probes.empty()) {
        // Hash probes are dependent random reads: latency, not bandwidth,
        // bounds the scan. Few rows per lane and one register set; 8 waves
        // per CU (measured on the 60M-row probe of the join workload against
        // a direct map: 0.61 ms against 0.75 ms at 16 waves per CU).
        t.block = 256;
        t.pipeline = 0;
        t.sub = std::max(1, 8 / t.items);
        t.waves_per_cu = 8;
        if (plan.dense_probes) {
            // Rank-map probes are cache hits: the pipelined shape with twice
            // the rows per tile is the fastest measured (0.480 ms against
            // 0.492 ms on the join workload's probe).
            t.pipeline = 1;
            t.sub *= 2;
        }
    }
    return t;
}

// A fold the scan cannot re-associate, in the reference's order
// (src/pass/loopgen.cpp:253-262: s = pred(row) ? next(s, row) : s over rows
// [lo, end) ascending): one wave; its 64 lanes load 64 consecutive rows'
// columns at a time (coalesced), then every lane folds the 64 rows in order
// from lane broadcasts, the state held identically in every lane. Lane 0
// writes the final state into the result struct.
static KernelSource gen_seq_scan(const Plan& plan, const std::vector<std::vector<bool>>& nullable,
                                 bool det) {
    KernelSource ks;
    ks.entry = "rf_qs";
    ks.sequential = true;
    ks.det = det;
    ks.tune.block = 64;
    ks.tune.items = 1;
    ks.tune.sub = 1;
    ks.tune.waves_per_cu = 1;
    Gen g(plan, nullable, ks);
    const ReduceComp& rc = plan.reduces.at(0);
    for (size_t f = 0; f < rc.fields.size(); f++)
        ks.acc.push_back({0, (int)f, rc.fields[f].type, AggOp::ADD});
    int off = 0, maxal = 1;
    for (auto& a : ks.acc) {
        const int w = a.type.width();
        off = (off + w - 1) / w * w;
        ks.acc_offset.push_back(off);
        off += w;
        maxal = std::max(maxal, w);
    }
    ks.acc_bytes = (off + maxal - 1) / maxal * maxal;

    std::ostringstream init;
    Emitter ei(g, Mode::SEQ, init, "    ");
    std::vector<std::string> iv;
    for (auto& f : rc.fields) iv.push_back(ei.expr(f.init));
    std::ostringstream body;
    Emitter em(g, Mode::SEQ, body, "            ");
    em.ntmp = ei.ntmp;
    const std::string pred = rc.seq_pred ? em.expr(rc.seq_pred) : std::string("true");
    std::vector<std::string> nx;
    for (auto& e : rc.seq_next) nx.push_back(em.expr(e));

    std::ostringstream o;
    o << "// Generated sequential fold for reffine function '" << plan.func_name << "'.\n";
    o << "extern \"C\" __global__ __launch_bounds__(64) void rf_qs(RfScanArgs L) {\n"
         "    const int lane = (int)threadIdx.x;\n";
    o << init.str();
    for (size_t f = 0; f < rc.fields.size(); f++) {
        const std::string T = cty(rc.fields[f].type);
        o << "    " << T << " s" << f << " = (" << T << ")" << iv[f] << ";\n";
    }
    // Blocks of R rows through LDS: the lanes load a block coalesced into
    // registers (R/64 rows per lane and column), store it to LDS and issue
    // the next block's loads; then every lane runs the block's rows in order,
    // each row's values an LDS broadcast (same address in every lane), 16
    // rows per unrolled step. (Rows broadcast from their lanes with
    // readlane one at a time cost about 100 cycles a row; the wave fold of
    // the group joins measured 513 -> 74 us with this layout.)
    int row_bytes = 0;
    for (int sl : g.seq_cols) row_bytes += ks.cols[sl].type.width();
    row_bytes += 4 * (int)g.seq_valid.size();
    int R = 512;
    while (R > 64 && (size_t)R * (size_t)std::max(row_bytes, 1) > 32 * 1024) R /= 2;
    const std::string RS = std::to_string(R), DS = std::to_string(R / 64);
    for (int sl : g.seq_cols) {
        const std::string T = cty(ks.cols[sl].type), S = std::to_string(sl);
        o << "    __shared__ " << T << " b" << S << "[" << RS << "];\n";
        o << "    " << T << " rl" << S << "[" << DS << "];\n";
    }
    for (int sl : g.seq_valid) {
        const std::string S = std::to_string(sl);
        o << "    __shared__ int bv" << S << "[" << RS << "];\n";
        o << "    int rv" << S << "[" << DS << "];\n";
    }
    auto loadR = [&](const std::string& at, const std::string& ind) {
        o << ind << "#pragma unroll\n" << ind << "for (int d = 0; d < " << DS << "; d++) {\n"
          << ind << "    const int64_t rl = " << at << " + d * 64 + lane;\n"
          << ind << "    const bool inl = rl < L.end;\n";
        for (int sl : g.seq_cols) {
            const std::string T = cty(ks.cols[sl].type), S = std::to_string(sl);
            o << ind << "    rl" << S << "[d] = inl ? rf_g((const " << T << "*)L.col[" << S
              << "])[rl] : (" << T << ")0;\n";
        }
        for (int sl : g.seq_valid) {
            const std::string S = std::to_string(sl);
            o << ind << "    rv" << S << "[d] = inl ? (int)rf_valid_bit(L.valid[" << S << "], rl) : 0;\n";
        }
        o << ind << "}\n";
    };
    loadR("L.lo", "    ");
    o << "    for (int64_t r0 = L.lo; r0 < L.end; r0 += " << RS << ") {\n";
    o << "        #pragma unroll\n        for (int d = 0; d < " << DS << "; d++) {\n";
    for (int sl : g.seq_cols) o << "            b" << sl << "[d * 64 + lane] = rl" << sl << "[d];\n";
    for (int sl : g.seq_valid) o << "            bv" << sl << "[d * 64 + lane] = rv" << sl << "[d];\n";
    o << "        }\n";
    loadR("r0 + " + RS, "        ");
    o << "        __builtin_amdgcn_wave_barrier();\n"
         "        const int n = (int)(L.end - r0 < " << RS << " ? L.end - r0 : " << RS << ");\n"
         "        #pragma unroll 16\n"
         "        for (int k = 0; k < n; k++) {\n"
         "            const int64_t row = r0 + k;\n"
         "            (void)row;\n";
    for (int sl : g.seq_cols) {
        const std::string T = cty(ks.cols[sl].type), S = std::to_string(sl);
        o << "            const " << T << " c" << S << " = b" << S << "[k];\n";
    }
    for (int sl : g.seq_valid) {
        const std::string S = std::to_string(sl);
        o << "            const bool v" << S << " = bv" << S << "[k] != 0;\n";
    }
    o << body.str();
    o << "            if (" << pred << ") {\n";
    for (size_t f = 0; f < nx.size(); f++) {
        const std::string T = cty(rc.fields[f].type);
        o << "                const " << T << " n" << f << " = (" << T << ")" << nx[f] << ";\n";
    }
    for (size_t f = 0; f < nx.size(); f++) o << "                s" << f << " = n" << f << ";\n";
    o << "            }\n        }\n        __builtin_amdgcn_wave_barrier();\n    }\n";
    o << "    if (lane == 0) {\n";
    for (size_t f = 0; f < rc.fields.size(); f++) {
        const std::string T = cty(rc.fields[f].type);
        o << "        *rf_gw((" << T << "*)((char*)L.result + " << ks.acc_offset[f] << ")) = s" << f
          << ";\n";
    }
    o << "    }\n}\n";
    ks.code = o.str();
    return ks;
}

// Union merge (an Op over _in(t, A) | _in(t, B)). Merge position m is the
// m-th step of the merge of A's and B's keys that takes A's key first on a
// tie; the state (i, j) before the step is the reference loop's state
// (LoopGen::build_loop, loopgen.cpp:44-110, over UnionSpace,
// iter_space.cpp:216-316: t = min of the two current keys, each side
// advances when its key is <= the other's) except when A's previous key
// equals B's current one: the reference advanced both sides at once there,
// the merge takes B's key as a step of its own, which emits nothing.
//
// Block b owns merge positions [b * TILE, (b + 1) * TILE): the A rows
// [splits[b], splits[b + 1]) and the B rows in between (merge-path splits,
// rf_merge_splits). It copies both key ranges, with the key before each,
// into LDS (coalesced; padded one slot in 8 so threads RF_MERGE_IT keys
// apart hit different banks), and each thread finds its own start in LDS and
// walks RF_MERGE_IT positions evaluating the predicate. rf_qu_count checks
// the keys strictly ascending (its range against the key before it; a split
// outside the tile's bounds also means unsorted keys: counters[0] is set)
// and writes the block's row count to partials[b] (0 for the extra block
// b = tiles). rf_qu_emit, after the exclusive scan of those counts, walks
// again, places the block's emitting states (tile-local rows and side taken)
// in LDS in merge order, then evaluates the outputs of row r of the block
// in thread r % BLOCK, so consecutive threads write consecutive rows.
static KernelSource gen_merge(const Plan& plan, const std::vector<std::vector<bool>>& nullable) {
    if (plan.ops.size() != 1 || !plan.reduces.empty())
        throw IRError("a union merge plan holds one Op");
    KernelSource ks;
    ks.entry = "rf_qu_count";
    ks.has_sel = true;
    ks.tune.block = RF_MERGE_BLOCK;
    ks.tune.items = RF_MERGE_IT;
    ks.tune.sub = 1;
    Gen g(plan, nullable, ks);
    const OpComp& op = plan.ops[0];
    const int ka = g.col_slot(plan.merge_a, 0, false), kb = g.col_slot(plan.merge_b, 0, false);
    ParamSlot pa{ParamSlot::LEN, types::IDX_T, 0, plan.merge_a, 0};
    ParamSlot pb{ParamSlot::LEN, types::IDX_T, 0, plan.merge_b, 0};
    const std::string na = g.param(pa, "$len." + std::to_string(plan.merge_a) + ".0");
    const std::string nb = g.param(pb, "$len." + std::to_string(plan.merge_b) + ".0");
    std::ostringstream pbody;
    Emitter ep(g, Mode::MERGE, pbody, "                ");
    const std::string pred = ep.expr(op.pred);
    std::ostringstream vbody;
    Emitter ev(g, Mode::MERGE, vbody, "            ");
    ev.ntmp = ep.ntmp;
    std::vector<std::string> vals;
    for (auto& v : op.values) vals.push_back(ev.expr(v));
    if (vals.size() > RF_MAXOUT) throw Unsupported("more than 16 output columns");
    ks.out_types = op.types;

    // A shifted side's keys: key + offset in wrapping int64 arithmetic.
    auto shifted = [](const std::string& load, int64_t off) {
        if (off == 0) return load;
        return "(int64_t)((uint64_t)" + load + " + " + std::to_string((uint64_t)off) + "ull)";
    };
    std::ostringstream o;
    o << "// Generated union merge for reffine function '" << plan.func_name << "'.\n"
         "#define RF_MP(x) ((x) + ((x) >> 3))\n"
         "#define RF_MERGE_LDS (RF_MP(RF_MERGE_TILE) + 1)\n"
         "template <bool EMIT>\n"
         "__device__ __forceinline__ void rf_qu_body(const RfScanArgs& L) {\n"
         "    __shared__ int64_t sA[RF_MERGE_LDS], sB[RF_MERGE_LDS];\n"
         "    __shared__ uint32_t srec[EMIT ? RF_MERGE_TILE : 1];\n"
         "    __shared__ uint64_t swave[RF_MERGE_BLOCK / 64];\n"
      << "    const int64_t na = " << na << ", nb = " << nb << ";\n"
      << "    const int64_t* __restrict__ KA = (const int64_t*)L.col[" << ka << "];\n"
      << "    const int64_t* __restrict__ KB = (const int64_t*)L.col[" << kb << "];\n"
      << "    const int64_t* __restrict__ splits = (const int64_t*)L.wave_dst;\n"
         "    uint32_t* __restrict__ bcnt = (uint32_t*)L.partials;\n"
         "    const int64_t M = na + nb;\n"
         "    const int64_t blk = blockIdx.x;\n"
         "    const int64_t d0 = blk * RF_MERGE_TILE;\n"
         "    if (d0 >= M) {\n"
         "        if (!EMIT && threadIdx.x == 0) bcnt[blk] = 0u;\n"
         "        return;\n"
         "    }\n"
         "    const int64_t d1 = d0 + RF_MERGE_TILE < M ? d0 + RF_MERGE_TILE : M;\n"
         "    const int len = (int)(d1 - d0);\n"
         "    const int64_t a0 = splits[blk];\n"
         "    int64_t a1 = splits[blk + 1];\n"
         "    bool bad = false;\n"
         "    if (a1 < a0) { a1 = a0; bad = true; }\n"
         "    if (a1 > a0 + len) { a1 = a0 + len; bad = true; }\n"
         "    const int64_t b0 = d0 - a0;\n"
         "    const int nA = (int)(a1 - a0), nB = len - nA;\n"
         "    // LDS slot x + 1 of a side holds its key at tile row x (slot 0: the\n"
         "    // key before the tile, 0 at row -1).\n"
         "    for (int x = threadIdx.x; x <= nA; x += RF_MERGE_BLOCK) {\n"
         "        const int64_t r = a0 - 1 + x;\n"
      << "        sA[RF_MP(x)] = r >= 0 ? " << shifted("rf_g(KA)[r]", plan.merge_off[0])
      << " : 0;\n"
         "    }\n"
         "    for (int x = threadIdx.x; x <= nB; x += RF_MERGE_BLOCK) {\n"
         "        const int64_t r = b0 - 1 + x;\n"
      << "        sB[RF_MP(x)] = r >= 0 ? " << shifted("rf_g(KB)[r]", plan.merge_off[1])
      << " : 0;\n"
         "    }\n"
         "    __syncthreads();\n"
         "    if (!EMIT) {\n"
         "        for (int x = threadIdx.x + 1; x <= nA; x += RF_MERGE_BLOCK)\n"
         "            if (a0 - 2 + x >= 0 && !(sA[RF_MP(x - 1)] < sA[RF_MP(x)])) bad = true;\n"
         "        for (int x = threadIdx.x + 1; x <= nB; x += RF_MERGE_BLOCK)\n"
         "            if (b0 - 2 + x >= 0 && !(sB[RF_MP(x - 1)] < sB[RF_MP(x)])) bad = true;\n"
         "        if (bad) atomicOr(L.counters, 1u);\n"
         "    }\n"
         "    const int q0 = (int)threadIdx.x * RF_MERGE_IT < len ? (int)threadIdx.x * RF_MERGE_IT : len;\n"
         "    int lo = q0 > nB ? q0 - nB : 0, hi = q0 < nA ? q0 : nA;\n"
         "    while (lo < hi) {\n"
         "        const int mid = (lo + hi) >> 1;\n"
         "        if (sA[RF_MP(mid + 1)] <= sB[RF_MP(q0 - mid)]) lo = mid + 1;\n"
         "        else hi = mid;\n"
         "    }\n"
         "    int ia = lo, jb = q0 - lo;\n"
         "    uint32_t rec[RF_MERGE_IT];\n"
         "    uint32_t emask = 0;\n"
         "#pragma unroll\n"
         "    for (int s = 0; s < RF_MERGE_IT; s++) {\n"
         "        rec[s] = 0u;\n"
         "        if (q0 + s < len) {\n"
         "            // ia + jb < nA + nB: one side at least has a row left in the tile.\n"
         "            const bool ain = ia < nA, bin = jb < nB;\n"
         "            const int64_t ka = ain ? sA[RF_MP(ia + 1)] : 0;\n"
         "            const int64_t kb = bin ? sB[RF_MP(jb + 1)] : 0;\n"
         "            const bool take0 = ain && (!bin || ka <= kb);\n"
         "            const int64_t i = a0 + ia, j = b0 + jb;\n"
         "            if (!(bin && i > 0 && sA[RF_MP(ia)] == kb)) {\n"
         "                const bool ok0 = i < na, ok1 = j < nb;\n"
         "                const int64_t key = take0 ? ka : kb;\n"
         "                (void)ok0; (void)ok1; (void)key;\n";
    o << pbody.str();
    o << "                if (" << pred << ") {\n"
         "                    rec[s] = (uint32_t)ia | ((uint32_t)jb << 12) | ((uint32_t)take0 << 24);\n"
         "                    emask |= 1u << s;\n"
         "                }\n"
         "            }\n"
         "            if (take0) ia++;\n"
         "            else jb++;\n"
         "        }\n"
         "    }\n"
         "    uint64_t total = 0;\n"
         "    const uint32_t base = (uint32_t)rf_block_excl_scan<RF_MERGE_BLOCK>(\n"
         "        (uint64_t)__builtin_popcount(emask), swave, total);\n"
         "    if (!EMIT) {\n"
         "        if (threadIdx.x == 0) bcnt[blk] = (uint32_t)total;\n"
         "        return;\n"
         "    }\n"
         "#pragma unroll\n"
         "    for (int s = 0; s < RF_MERGE_IT; s++)\n"
         "        if ((emask >> s) & 1u)\n"
         "            srec[base + __builtin_popcount(emask & ((1u << s) - 1u))] = rec[s];\n"
         "    __syncthreads();\n"
         "    const int64_t off = bcnt[blk];\n"
         "    for (int r = threadIdx.x; r < (int)total; r += RF_MERGE_BLOCK) {\n"
         "        const uint32_t e = srec[r];\n"
         "        const int ia = (int)(e & 4095u), jb = (int)((e >> 12) & 4095u);\n"
         "        const bool take0 = ((e >> 24) & 1u) != 0u;\n"
         "        const int64_t i = a0 + ia, j = b0 + jb;\n"
         "        const bool ok0 = i < na, ok1 = j < nb;\n"
         "        const int64_t key = take0 ? sA[RF_MP(ia + 1)] : sB[RF_MP(jb + 1)];\n"
         "        (void)ok0; (void)ok1; (void)key;\n";
    o << vbody.str();
    for (size_t k = 0; k < vals.size(); k++) {
        const std::string T = cty(op.types[k]);
        o << "            ((" << T << "*)L.out[" << k << "])[off + r] = (" << T << ")" << vals[k]
          << ";\n";
    }
    o << "    }\n"
         "}\n\n"
         "extern \"C\" __global__ __launch_bounds__(RF_MERGE_BLOCK) void rf_qu_count(RfScanArgs L) { "
         "rf_qu_body<false>(L); }\n"
         "extern \"C\" __global__ __launch_bounds__(RF_MERGE_BLOCK) void rf_qu_emit(RfScanArgs L) { "
         "rf_qu_body<true>(L); }\n";
    ks.code = o.str();
    return ks;
}

KernelSource gen_scan(const Plan& plan, const std::vector<std::vector<bool>>& nullable,
                      const Tuning& tune, bool det, const std::vector<bool>& pack_probes) {
    if (plan.source == Plan::MERGE) return gen_merge(plan, nullable);
    if (plan.sequential) return gen_seq_scan(plan, nullable, det);
    KernelSource ks;
    ks.tune = tune;
    ks.det = det;
    Gen g(plan, nullable, ks);
    g.pack_probes = pack_probes;
    if (tune.block % 64 != 0 || tune.block > 1024) throw IRError("block size");
    if (16 % tune.items != 0) throw IRError("ITEMS must divide 16");
    ks.has_sel = !plan.ops.empty();
    // Load group: the largest divisor of SUB within the register budget of
    // one set of loads (a wave-tile of SUB sub-steps is evaluated in SUB/LSUB
    // groups; the rows each lane folds, and their order, do not change).
    const int lcap = load_cap(plan, tune.items, tune.pipeline != 0);
    int lsub = 1;
    for (int d = 1; d <= tune.sub; d++)
        if (tune.sub % d == 0 && d <= lcap) lsub = d;
    if (tune.lsub > 0) {
        if (tune.sub % tune.lsub != 0) throw IRError("LSUB must divide SUB");
        lsub = tune.lsub;
    }

    // Accumulator layout: one field per updated state field.
    for (size_t c = 0; c < plan.reduces.size(); c++)
        for (size_t f = 0; f < plan.reduces[c].fields.size(); f++) {
            const Field& fl = plan.reduces[c].fields[f];
            if (!fl.updated) continue;
            ks.acc.push_back({(int)c, (int)f, fl.type, fl.op});
        }
    // Struct layout with natural alignment.
    int off = 0, maxal = 1;
    for (auto& a : ks.acc) {
        int w = a.type.width();
        off = (off + w - 1) / w * w;
        ks.acc_offset.push_back(off);
        off += w;
        maxal = std::max(maxal, w);
    }
    ks.acc_bytes = ks.acc.empty() ? 0 : (off + maxal - 1) / maxal * maxal;

    // ---- rows(): row code
    std::ostringstream body;
    Emitter em(g, Mode::ROWS, body, "            ");
    std::vector<std::string> acc_pred, acc_contrib;
    for (auto& a : ks.acc) {
        const Field& fl = plan.reduces[a.comp].fields[a.field];
        acc_pred.push_back(em.expr(fl.pred));
        acc_contrib.push_back(em.expr(fl.contrib));
    }
    std::string sel_pred;
    if (ks.has_sel) sel_pred = em.expr(plan.ops[0].pred);
    // Early materialisation: the outputs from the loaded rows, for every row
    // (every leaf is safe to evaluate on a row that is not selected: probed
    // columns read only on a hit).
    const bool early = ks.has_sel && plan.early_sel;
    std::vector<std::string> early_vals;
    if (early) {
        for (auto& v : plan.ops[0].values) early_vals.push_back(em.expr(v));
        if (early_vals.size() > 16) throw Unsupported("more than 16 output columns");
    }
    g.carried = em.probes_done;

    std::ostringstream o;
    o << "// Generated row code for reffine function '" << plan.func_name << "'.\n";
    o << "struct Q {\n";
    o << "    enum { BLOCK = " << tune.block << ", ITEMS = " << tune.items
      << ", SUB = " << tune.sub << ", LSUB = " << lsub << ", NACC = " << ks.acc.size()
      << ", HAS_SEL = " << (ks.has_sel && !plan.lds_agg ? 1 : 0) << ", DET = " << (det ? 1 : 0)
      << ", PIPE = " << (tune.pipeline ? 1 : 0) << " };\n";
    o << "    struct Acc {";
    for (size_t i = 0; i < ks.acc.size(); i++) o << " " << cty(ks.acc[i].type) << " a" << i << ";";
    if (ks.acc.empty()) o << " int unused;";
    o << " };\n";
    o << "    struct Carry {";
    for (int j : g.carried) o << " int64_t p" << j << "[LSUB][ITEMS];";
    for (size_t j = 0; j < early_vals.size(); j++)
        o << " " << cty(plan.ops[0].types[j]) << " o" << j << "[LSUB][ITEMS];";
    if (g.carried.empty() && early_vals.empty()) o << " int unused;";
    o << " };\n";
    o << "    static __device__ __forceinline__ void acc_zero(Acc& a) {";
    for (size_t i = 0; i < ks.acc.size(); i++)
        o << " a.a" << i << " = " << fold_identity(ks.acc[i].op, ks.acc[i].type) << ";";
    o << " }\n";
    o << "    static __device__ __forceinline__ void acc_combine(Acc& a, const Acc& b) {";
    for (size_t i = 0; i < ks.acc.size(); i++) {
        std::string ai = "a.a" + std::to_string(i), bi = "b.a" + std::to_string(i);
        o << " " << ai << " = " << combine_expr(ks.acc[i].op, ks.acc[i].type, ai, bi) << ";";
    }
    o << " }\n";
    o << "    static __device__ __forceinline__ void acc_shfl_xor(Acc& a, int o) {\n"
      << "        Acc b;";
    for (size_t i = 0; i < ks.acc.size(); i++)
        o << " b.a" << i << " = rf_shfl_xor(a.a" << i << ", o);";
    o << "\n        acc_combine(a, b);\n    }\n";    // XOR: the butterfly level a mirror control stands for (0 for the quad
    // permutations, which are that xor). A float min/max field is not
    // commutative once a NaN is involved (lanes of a quad may then differ), so
    // it moves by the xor shuffle itself at the mirror levels.
    o << "    template <int CTRL, int XOR>\n"
      << "    static __device__ __forceinline__ void acc_dpp(Acc& a) {\n"
      << "        Acc b;";
    for (size_t i = 0; i < ks.acc.size(); i++) {
        const bool exact = ks.acc[i].type.is_float() &&
                           (ks.acc[i].op == AggOp::MIN || ks.acc[i].op == AggOp::MAX);
        o << " b.a" << i << " = " << (exact ? "rf_dpp_x<CTRL, XOR>" : "rf_dpp<CTRL>") << "(a.a" << i
          << ");";
    }
    o << "\n        acc_combine(a, b);\n    }\n";

    // Vector-loaded columns become parameters of row_block.
    std::string vparams, vargs, vargs_s;
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& c = ks.cols[s];
        if (!c.vector) continue;
        const std::string T = cty(c.type), S = std::to_string(s);
        vparams += ", const RfVec<" + T + ", ITEMS>& c" + S;
        vargs += ", c" + S;
        vargs_s += ", c" + S + "[s]";
        if (c.nullable) {
            vparams += ", uint32_t vb" + S;
            vargs += ", vb" + S;
            vargs_s += ", vb" + S + "[s]";
        }
    }
    // rows(): one sub-step with bounds (edge tiles).
    o << "    template <bool FULL>\n"
      << "    static __device__ __forceinline__ uint32_t rows(Acc& a, const RfScanArgs& L, int64_t r, "
         "Carry& cy, int s) {\n";
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& c = ks.cols[s];
        if (!c.vector) continue;
        const std::string T = cty(c.type);
        o << "        const RfVec<" << T << ", ITEMS> c" << s << " = FULL ? rf_ldv<" << T
          << ", ITEMS>((const " << T << "*)L.col[" << s << "] + r) : rf_ldv_tail<" << T
          << ", ITEMS>((const " << T << "*)L.col[" << s << "], r, L.end);\n";
        if (c.nullable)
            o << "        const uint32_t vb" << s << " = rf_valid_bits<ITEMS>(L.valid[" << s
              << "], r);\n";
    }
    o << "        return row_block<FULL>(a, L, r, cy, s" << vargs << ");\n    }\n";
    // Tile: the registers of one interior wave-tile; load() issues every
    // sub-step's loads, compute() evaluates the rows (the skeleton overlaps
    // the next tile's load() with this tile's compute()).
    o << "    struct Tile {";
    bool any_vec = false;
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& c = ks.cols[s];
        if (!c.vector) continue;
        any_vec = true;
        o << " RfVec<" << cty(c.type) << ", ITEMS> c" << s << "[LSUB];";
        if (c.nullable) o << " uint32_t vb" << s << "[LSUB];";
    }
    if (!any_vec) o << " int unused;";
    o << " };\n";
    o << "    static __device__ __forceinline__ void load(Tile& T, const RfScanArgs& L, "
         "int64_t row0, uint32_t lane) {\n"
      << "        const int64_t rb = row0 + (int64_t)lane * ITEMS;\n        (void)rb; (void)T;\n"
      << "#pragma unroll\n        for (int s = 0; s < LSUB; s++) {\n"
      << "            const int64_t r = rb + (int64_t)s * 64 * ITEMS;\n            (void)r;\n";
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& c = ks.cols[s];
        if (!c.vector) continue;
        const std::string T = cty(c.type);
        o << "            T.c" << s << "[s] = rf_ldv<" << T << ", ITEMS>((const " << T << "*)L.col["
          << s << "] + r);\n";
        if (c.nullable)
            o << "            T.vb" << s << "[s] = rf_valid_bits<ITEMS>(L.valid[" << s << "], r);\n";
    }
    o << "        }\n    }\n";
    std::string targs_s;
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& c = ks.cols[s];
        if (!c.vector) continue;
        targs_s += ", T.c" + std::to_string(s) + "[s]";
        if (c.nullable) targs_s += ", T.vb" + std::to_string(s) + "[s]";
    }
    o << "    static __device__ __forceinline__ void compute(Acc& a, const Tile& T, "
         "const RfScanArgs& L, int64_t row0, uint32_t lane, uint32_t* bits, Carry& cy) {\n"
      << "        const int64_t rb = row0 + (int64_t)lane * ITEMS;\n        (void)T;\n"
      << "#pragma unroll\n        for (int s = 0; s < LSUB; s++) {\n"
      << "            const int64_t r = rb + (int64_t)s * 64 * ITEMS;\n"
      << "            bits[s] = row_block<true>(a, L, r, cy, s" << targs_s << ");\n        }\n    }\n";
    // row_block(): the query's row code over ITEMS loaded rows.
    o << "    template <bool FULL>\n"
      << "    static __device__ __forceinline__ uint32_t row_block(Acc& a, const RfScanArgs& L, "
         "int64_t r, Carry& cy, int s" << vparams << ") {\n";
    o << "        uint32_t sel = 0;\n        (void)cy; (void)s;\n";
    o << "#pragma unroll\n        for (int k = 0; k < ITEMS; k++) {\n";
    o << "            const int64_t row = r + k;\n";
    o << "            const bool inb = FULL || (row >= L.lo && row < L.end);\n";
    o << "            (void)row;\n";
    o << body.str();
    // Accumulators sharing a predicate update under one branch: lanes whose
    // row fails it are masked off instead of selecting each accumulator's
    // old value (two 32-bit selects per 64-bit accumulator).
    {
        std::vector<bool> done(ks.acc.size(), false);
        for (size_t i = 0; i < ks.acc.size(); i++) {
            if (done[i]) continue;
            std::vector<size_t> same;
            for (size_t j = i; j < ks.acc.size(); j++)
                if (!done[j] && acc_pred[j] == acc_pred[i]) same.push_back(j);
            if (same.size() == 1) {
                std::string ai = "a.a" + std::to_string(i);
                o << "            " << ai << " = (inb && " << acc_pred[i] << ") ? "
                  << combine_expr(ks.acc[i].op, ks.acc[i].type, ai, acc_contrib[i]) << " : " << ai
                  << ";\n";
                done[i] = true;
                continue;
            }
            o << "            if (inb && " << acc_pred[i] << ") {\n";
            for (size_t j : same) {
                std::string aj = "a.a" + std::to_string(j);
                o << "                " << aj << " = "
                  << combine_expr(ks.acc[j].op, ks.acc[j].type, aj, acc_contrib[j]) << ";\n";
                done[j] = true;
            }
            o << "            }\n";
        }
    }
    if (ks.has_sel) o << "            sel |= (uint32_t)(inb && " << sel_pred << ") << k;\n";
    for (size_t j = 0; j < early_vals.size(); j++) {
        const std::string T = cty(plan.ops[0].types[j]);
        o << "            cy.o" << j << "[s][k] = (" << T << ")" << early_vals[j] << ";\n";
    }
    o << "        }\n        return sel;\n    }\n";

    // ---- emit(): stages a qualifying row's index (and the probe results its
    // row code computed) -- no loads in the scan; move(): the placement
    // kernel reads the row's outputs by index and writes them at their
    // output position (late materialisation: the scan never waits on a
    // selected row's reads).
    std::ostringstream mv;
    // The staged columns: (C type, value at (s, k)) in stage order.
    std::vector<std::pair<std::string, std::string>> stage_vals;
    if (early) {
        // The staged values are the outputs; the placement kernel copies.
        for (size_t j = 0; j < early_vals.size(); j++) {
            const std::string T = cty(plan.ops[0].types[j]), J = std::to_string(j);
            stage_vals.push_back({T, "cy.o" + J + "[s][k]"});
            mv << "        rf_gw((" << T << "*)L.out[" << J << "])[dst] = rf_g((const " << T
               << "*)L.stage[" << J << "])[src];\n";
            ks.stage_types.push_back(plan.ops[0].types[j]);
            ks.out_types.push_back(plan.ops[0].types[j]);
        }
    } else if (ks.has_sel) {
        std::ostringstream mb;
        Emitter me(g, Mode::MOVE, mb, "        ");
        std::vector<std::string> vals;
        for (auto& v : plan.ops[0].values) vals.push_back(me.expr(v));
        if (vals.size() > 16) throw Unsupported("more than 16 output columns");
        ks.stage_types.push_back(types::I64);  // row index
        stage_vals.push_back({"int64_t", "row"});
        mv << "        const int64_t row = rf_g((const int64_t*)L.stage[0])[src];\n";
        for (int j : me.staged) {
            const std::string S = std::to_string(ks.stage_types.size());
            stage_vals.push_back({"int64_t", "cy.p" + std::to_string(j) + "[s][k]"});
            mv << "        const int64_t sc" << j << " = rf_g((const int64_t*)L.stage[" << S
               << "])[src];\n";
            ks.stage_types.push_back(types::I64);
        }
        if (ks.stage_types.size() > 16) throw Unsupported("more than 16 staged values");
        mv << mb.str();
        for (size_t j = 0; j < vals.size(); j++) {
            const std::string T = cty(plan.ops[0].types[j]), J = std::to_string(j);
            mv << "        rf_gw((" << T << "*)L.out[" << J << "])[dst] = (" << T << ")" << vals[j]
               << ";\n";
            ks.out_types.push_back(plan.ops[0].types[j]);
        }
    } else {
        mv << "        (void)L; (void)src; (void)dst;\n";
    }
    // A group join's probe scan (joined rows are a large share of the probe
    // rows, staged with their values) compacts a load group's rows in the
    // wave's LDS buffer and writes them to the staging region with one
    // coalesced store per 64 rows and column (rf_tile_emit_lds), when the
    // buffers fit: per-row stores leave most lanes of each store idle.
    // Measured on the 60M-row join probe (16 % of rows staged, 12 bytes
    // each): 0.478 -> 0.451 ms. A sparse selection keeps per-row stores (the
    // C2 scan, 0.13 % selected: 0.192 ms against 0.218 ms through LDS, the
    // larger kernel costing more than the stores it saves).
    int64_t row_bytes = 0;
    for (auto& t : ks.stage_types) row_bytes += t.width();
    const int64_t group_rows = (int64_t)lsub * 64 * tune.items;
    const int64_t block_lds = (int64_t)(tune.block / 64) * group_rows * row_bytes;
    const int64_t blocks_per_cu = std::max(1, tune.waves_per_cu / std::max(1, tune.block / 64));
    const bool lds_emit = early && !plan.lds_agg && !stage_vals.empty() && block_lds <= 48 * 1024 &&
                          blocks_per_cu * block_lds <= 96 * 1024;
    std::ostringstream es, el, fl, eb;
    for (size_t j = 0; j < stage_vals.size(); j++) {
        const std::string& T = stage_vals[j].first;
        const std::string J = std::to_string(j);
        es << "        rf_gw((" << T << "*)L.stage[" << J << "])[pos] = " << stage_vals[j].second
           << ";\n";
        el << "        b.c" << J << "[i] = " << stage_vals[j].second << ";\n";
        fl << "        rf_gw((" << T << "*)L.stage[" << J << "])[pos] = b.c" << J << "[i];\n";
        eb << " " << T << " c" << J << "[LSUB * 64 * ITEMS];";
    }
    if (stage_vals.empty()) es << "        (void)L; (void)row; (void)pos;\n";
    o << "    static __device__ __forceinline__ void emit(const RfScanArgs& L, int64_t row, int64_t pos, "
         "const Carry& cy, int s, int k) {\n        (void)cy; (void)s; (void)k;\n"
      << es.str() << "    }\n";
    o << "    enum { LDS_EMIT = " << (lds_emit ? 1 : 0) << " };\n";
    ks.lds_emit = lds_emit;
    o << "    struct EmitBuf {" << (lds_emit ? eb.str() : std::string(" int unused;")) << " };\n";
    o << "    static __device__ __forceinline__ void emit_lds(EmitBuf& b, int i, int64_t row, "
         "const Carry& cy, int s, int k) {\n        (void)b; (void)i; (void)row; (void)cy; (void)s; "
         "(void)k;\n"
      << (lds_emit ? el.str() : std::string()) << "    }\n";
    // The partition key of a staged row (a group join's rank, staged first).
    o << "    static __device__ __forceinline__ int32_t part_key(const EmitBuf& b, int i) {\n"
         "        (void)b; (void)i;\n"
      << (lds_emit && early ? "        return (int32_t)b.c0[i];\n" : "        return 0;\n") << "    }\n";
    o << "    static __device__ __forceinline__ void flush(const RfScanArgs& L, const EmitBuf& b, int i, "
         "int64_t pos) {\n        (void)L; (void)b; (void)i; (void)pos;\n"
      << (lds_emit ? fl.str() : std::string()) << "    }\n";
    // Pre-aggregation in LDS (a group join's probe, plan.lds_agg): each
    // joined row's contributions go straight into its block's LDS slots --
    // slot f*gpp + rank for updated field f -- with LDS atomics; no row is
    // staged. Fields as rf_gj_agg takes them (gj_agg_ok).
    int nslot = 0;
    std::ostringstream ag, ai;
    if (plan.lds_agg) {
        const GroupJoinPlan& gj = *plan.lds_agg;
        if (!early || !gj_agg_ok(gj)) throw IRError("LDS pre-aggregation needs an agg-able group join");
        ag << "        const int64_t g = (int64_t)cy.o0[s][k];\n";
        for (const GField& gf : gj.fields) {
            if (!gf.updated) continue;
            const int sl = nslot++;
            const std::string T = cty(gf.type);
            const std::string v = gf.const_contrib ? "(" + T + ")" + g.const_param(gf.const_contrib)
                                                   : "(" + T + ")cy.o" + std::to_string(1 + gf.pay) +
                                                         "[s][k]";
            const std::string a = "acc[" + std::to_string(sl) + " * gpp + g]";
            std::string upd;
            if (gf.op == AggOp::ADD && gf.type.bt == BT::FLOAT64)
                upd = "atomicAdd((double*)&" + a + ", (double)" + v + ")";
            else if (gf.op == AggOp::ADD && gf.type.bt == BT::FLOAT32)
                upd = "atomicAdd((float*)&" + a + ", (float)" + v + ")";
            else if (gf.op == AggOp::ADD)
                upd = "atomicAdd(&" + a + ", (unsigned long long)(long long)" + v + ")";
            else {
                const char* fn = gf.op == AggOp::MIN ? "atomicMin" : "atomicMax";
                upd = gf.type.is_unsigned()
                          ? std::string(fn) + "(&" + a + ", (unsigned long long)" + v + ")"
                          : std::string(fn) + "((long long*)&" + a + ", (long long)" + v + ")";
            }
            if (gf.pred_pay >= 0)
                ag << "        if (cy.o" << 1 + gf.pred_pay << "[s][k]) " << upd << ";\n";
            else
                ag << "        " << upd << ";\n";
            ai << "        case " << sl << ": return " << gj_slot_identity(gf) << ";\n";
        }
    }
    o << "    enum { LDS_AGG = " << (plan.lds_agg ? 1 : 0) << ", NSLOT = " << nslot << " };\n";
    o << "    static __device__ __forceinline__ unsigned long long agg_ident(int slot) {\n"
         "        switch (slot) {\n" << ai.str() << "        default: return 0ull;\n        }\n    }\n";
    o << "    static __device__ __forceinline__ void agg_row(const RfScanArgs& L, unsigned long long* acc, "
         "int64_t gpp, const Carry& cy, int s, int k) {\n        (void)L; (void)acc; (void)gpp; (void)cy; "
         "(void)s; (void)k;\n" << ag.str() << "    }\n";
    o << "    static __device__ __forceinline__ void move(const RfScanArgs& L, int64_t src, int64_t dst) {\n"
      << mv.str() << "    }\n};\n\n";
    if (plan.lds_agg) {
        // Nothing is staged or placed: the block slots are the output.
        ks.has_sel = false;
        ks.stage_types.clear();
        ks.out_types.clear();
        ks.lds_agg_slots = nslot;
    }
    o << "extern \"C\" __global__ __launch_bounds__(" << tune.block
      << ") void rf_q(RfScanArgs L) { rf_scan<Q>(L); }\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_q_gather(RfScanArgs L) { "
         "rf_gather<Q>(L); }\n";
    o << "extern \"C\" __global__ __launch_bounds__(" << tune.block
      << ") void rf_qb(RfBatchArgs B) { rf_batch_scan<Q>(B); }\n";
    o << "extern \"C\" __global__ __launch_bounds__(" << tune.block
      << ") void rf_qb_done(RfBatchArgs B) { rf_batch_done<Q>(B); }\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_qb_gather(RfBatchArgs B) { "
         "rf_batch_gather<Q>(B); }\n";
    ks.code = o.str();
    return ks;
}

KernelSource gen_seg_fold(const SegReducePlan& sp, const Plan& plan,
                          const std::vector<std::vector<bool>>& nullable,
                          const std::vector<bool>& pack_probes) {
    (void)plan;
    KernelSource ks;
    ks.entry = "rf_seg_fold";
    ks.tune.block = 256;
    if (sp.sub.probes.size() > 4) throw Unsupported("more than 4 probed tables");
    Gen g(sp.sub, nullable, ks);
    g.pack_probes = pack_probes;
    bool any_pack = false;
    for (bool b : pack_probes) any_pack = any_pack || b;
    // A block folds 256 consecutive runs, a thread per run, each run's rows
    // in row order (the reference's SubVector loop). The rows the block's
    // runs cover are first copied to LDS with coalesced loads (a thread's
    // own run would otherwise read a few rows at a stride of the neighbouring
    // runs' lengths); blocks whose runs span more rows than the LDS copy
    // holds read global memory.
    std::ostringstream init;
    Emitter ei(g, Mode::FINAL, init, "        ");
    // Tables read at the run's key: probed once per run, before its rows
    // (probes keyed by a row's columns stay in the row loop).
    auto row_free = [](const N& e) {
        std::function<bool(const N&)> f = [&](const N& x) -> bool {
            if (!x) return true;
            const std::string ln = leaf_name(x);
            if (ln.rfind("$col", 0) == 0 || ln.rfind("$valid", 0) == 0 || ln == "$row")
                return false;
            for (auto& a : x->args)
                if (!f(a)) return false;
            return true;
        };
        return f(e);
    };
    std::set<int> per_run;
    for (size_t j = 0; j < sp.sub.probes.size(); j++) {
        const Probe& pr = sp.sub.probes[j];
        if (row_free(pr.key) && row_free(pr.guard)) {
            ei.probe((int)j);
            per_run.insert((int)j);
        }
    }
    // So are those tables' columns the fold reads (per run, not per row).
    std::function<void(const N&)> pcols = [&](const N& e) {
        if (!e) return;
        if (e->kind == K::Sym && leaf_name(e).rfind("$pcol.", 0) == 0) {
            if (per_run.count(std::stoi(split(leaf_name(e), '.')[1]))) ei.expr(e);
            return;
        }
        for (auto& a : e->args) pcols(a);
    };
    for (auto& f : sp.fields) {
        if (!f.updated) continue;
        pcols(f.pred);
        pcols(f.contrib);
    }
    std::vector<std::string> iv;
    for (auto& f : sp.fields) iv.push_back(ei.expr(f.init));
    auto body = [&](bool lds, std::ostringstream& out) {
        std::ostringstream b;
        g.lds_rows = lds;
        Emitter eb(g, Mode::MOVE, b, "                ");
        eb.ntmp = ei.ntmp;
        eb.probes_done = ei.probes_done;
        eb.memo = ei.memo;
        std::vector<std::string> pr(sp.fields.size()), ct(sp.fields.size());
        for (size_t f = 0; f < sp.fields.size(); f++) {
            if (!sp.fields[f].updated) continue;
            pr[f] = eb.expr(sp.fields[f].pred);
            ct[f] = eb.expr(sp.fields[f].contrib);
        }
        g.lds_rows = false;
        // Branch-free probes: four rows unrolled issue their reads together
        // (the fold itself stays in row order).
        if (any_pack) out << "#pragma unroll 4\n";
        out << "            for (int64_t row = lo; row < hi; row++) {\n" << b.str();
        for (size_t f = 0; f < sp.fields.size(); f++) {
            if (!sp.fields[f].updated) continue;
            const std::string S = "s" + std::to_string(f);
            out << "                " << S << " = " << pr[f] << " ? "
                << combine_expr(sp.fields[f].op, sp.fields[f].type, S, ct[f]) << " : " << S
                << ";\n";
        }
        out << "            }\n";
        return (int)eb.ntmp;
    };
    std::ostringstream loop_lds, loop_glb;
    int nt = body(true, loop_lds);
    nt = std::max(nt, body(false, loop_glb));
    std::ostringstream ob;
    Emitter eo(g, Mode::FINAL, ob, "            ");
    eo.ntmp = nt;
    eo.probes_done = ei.probes_done;
    std::vector<std::string> ov;
    for (auto& v : sp.out_values) ov.push_back(eo.expr(v));
    if (ov.size() > 16) throw Unsupported("more than 16 run reduction outputs");
    // LDS rows per block: 6 per run (TPC-H orders hold 1..7 lines, 4 on
    // average), at most 40 KiB over the staged columns (4 blocks per CU).
    int row_bytes = 0;
    for (int sl : g.lds_slots) row_bytes += ks.cols[sl].type.width();
    int cap = 0, cap_long = 0;
    if (row_bytes > 0) {
        cap = std::min(1536, (40960 / row_bytes) / 256 * 256);
        cap_long = std::min(8192, (65536 / row_bytes) / 256 * 256);
    }
    const bool staged = cap >= 256 && !sp.ranged;
    std::string lo_e, hi_e;
    std::ostringstream bnd;
    if (sp.ranged) {
        // A ranged fold's rows [lo, hi): expressions of its key.
        Emitter er(g, Mode::FINAL, bnd, "            ");
        er.ntmp = nt + 1000;
        lo_e = er.expr(sp.lo);
        hi_e = er.expr(sp.hi);
    }

    for (size_t j = 0; j < ov.size(); j++) ks.out_types.push_back(sp.out_types[j]);
    auto emit_kernel = [&](std::ostringstream& o, const char* kname, int cap, bool staged) {
    o << "extern \"C\" __global__ __launch_bounds__(256) void " << kname << "(RfSegArgs L) {\n";
    if (staged) {
        o << "    constexpr int64_t CAP = " << cap << ";\n";
        for (int sl : g.lds_slots)
            o << "    __shared__ " << cty(ks.cols[sl].type) << " sm" << sl << "[CAP];\n";
    }
    o << "    for (int64_t base = blockIdx.x * (int64_t)256; base < L.runs;\n"
         "         base += (int64_t)gridDim.x * 256) {\n";
    o << "        const int64_t run = base + threadIdx.x;\n";
    // The run's key, its probes and the fold's init come first, so that
    // their loads overlap the copy below.
    o << "        const int64_t rr = run < L.runs ? run : L.runs - 1;\n";
    if (sp.ranged)
        o << "        const int64_t key = L.key0 + rr;\n        (void)key;\n";
    else
        o << "        const " << cty(sp.key_type) << " key = ((const " << cty(sp.key_type)
          << "*)L.keys)[rr];\n        (void)key;\n";
    o << init.str();
    if (staged) {
        o << "        const int64_t last = (base + 256 < L.runs ? base + 256 : L.runs) - 1;\n";
        o << "        int64_t blo = base > 0 ? (int64_t)L.run_ends[base - 1] : 0;\n";
        o << "        int64_t bhi = (int64_t)L.run_ends[last];\n";
        o << "        blo = blo < L.rows ? blo : L.rows;\n";
        o << "        bhi = bhi < L.rows ? bhi : L.rows;\n";
        o << "        bhi = bhi > blo ? bhi : blo;\n";
        o << "        const bool in_lds = bhi - blo <= CAP;\n";
        o << "        __syncthreads();  // the previous runs' LDS reads are done\n";
        // U rows per lane per pass, all loads issued before the LDS stores
        // (a store waits for its own load; one row a pass serialises them).
        // Up to 192 load bytes per lane: a block's ~1000 rows of TPC-H
        // lineitem (24-byte rows) in one pass; 0.345 -> 0.337 ms on Q3's
        // fold and 0.377 -> 0.363 ms on Q12's against 64 bytes (2 rows).
        const int U = std::max(2, std::min(8, 192 / row_bytes));
        o << "        if (in_lds) {\n";
        o << "            const int64_t nb = bhi - blo;\n";
        o << "            for (int64_t i0 = threadIdx.x; i0 < nb; i0 += 256 * " << U << ") {\n";
        for (int sl : g.lds_slots)
            o << "                " << cty(ks.cols[sl].type) << " r" << sl << "[" << U << "];\n";
        o << "#pragma unroll\n";
        o << "                for (int u = 0; u < " << U << "; u++) {\n";
        o << "                    const int64_t i = i0 + u * 256;\n";
        o << "                    if (i < nb) {\n";
        for (int sl : g.lds_slots) {
            const std::string T = cty(ks.cols[sl].type);
            o << "                        r" << sl << "[u] = rf_g((const " << T << "*)L.col["
              << sl << "])[blo + i];\n";
        }
        o << "                    }\n                }\n";
        o << "#pragma unroll\n";
        o << "                for (int u = 0; u < " << U << "; u++) {\n";
        o << "                    const int64_t i = i0 + u * 256;\n";
        o << "                    if (i < nb) {\n";
        for (int sl : g.lds_slots)
            o << "                        sm" << sl << "[i] = r" << sl << "[u];\n";
        o << "                    }\n                }\n";
        o << "            }\n";
        o << "        }\n";
        o << "        __syncthreads();\n";
    }
    o << "        if (run < L.runs) {\n";
    if (sp.ranged) {
        o << bnd.str();
        o << "            int64_t lo = (int64_t)" << lo_e << ";\n";
        o << "            int64_t hi = (int64_t)" << hi_e << ";\n";
        o << "            lo = lo < 0 ? 0 : lo;\n";
        o << "            hi = hi < L.rows ? hi : L.rows;\n";
    } else {
        o << "            int64_t lo = run > 0 ? (int64_t)L.run_ends[run - 1] : 0;\n";
        o << "            const int64_t hi0 = (int64_t)L.run_ends[run];\n";
        o << "            int64_t hi = hi0 < L.rows ? hi0 : L.rows;\n";
    }
    for (size_t f = 0; f < sp.fields.size(); f++) {
        const std::string T = cty(sp.fields[f].type);
        o << "            " << T << " s" << f << " = (" << T << ")" << iv[f] << ";\n";
    }
    if (staged) {
        // Rows outside the block's copy are never read (run ends that do
        // not ascend leave such runs short, not out of bounds).
        o << "            if (in_lds) {\n";
        o << "                lo = lo > blo ? lo : blo;\n";
        o << "                hi = hi < bhi ? hi : bhi;\n";
        o << loop_lds.str();
        o << "            } else {\n";
        o << loop_glb.str();
        o << "            }\n";
    } else {
        o << loop_glb.str();
    }
    o << ob.str();
    for (size_t j = 0; j < ov.size(); j++) {
        const std::string T = cty(sp.out_types[j]);
        o << "            ((" << T << "*)L.out[" << j << "])[run] = (" << T << ")" << ov[j] << ";\n";
    }
    o << "        }\n    }\n}\n";
    };
    std::ostringstream o;
    o << "// Generated run reduction for '" << sp.sym_name << "'.\n";
    emit_kernel(o, "rf_seg_fold", cap, staged);
    // For runs longer than 6 rows on average (graph adjacency runs): the same
    // fold with the copy sized to 64 KiB (at most 32 rows per run), chosen
    // at launch from the table's rows per run.
    if (staged && cap_long > cap) {
        emit_kernel(o, "rf_seg_fold_long", cap_long, true);
        ks.alt_cap = cap_long;
    }
    ks.cap = cap;
    // A wave per fold, for runs or windows of 32 rows and more (a window of
    // Nbody's pair table holds 2047): the 64 lanes load 64 rows at a time
    // (coalesced) and evaluate their predicates and contributions; every lane
    // then folds the 64 in row order from lane broadcasts -- the same
    // operations in the same order as one thread's loop.
    {
        std::ostringstream wb;
        g.lds_rows = false;
        Emitter ew(g, Mode::MOVE, wb, "                ");
        ew.ntmp = nt + 2000;
        ew.probes_done = ei.probes_done;
        ew.memo = ei.memo;
        std::vector<std::string> wp(sp.fields.size()), wc(sp.fields.size());
        for (size_t f = 0; f < sp.fields.size(); f++) {
            if (!sp.fields[f].updated) continue;
            wp[f] = ew.expr(sp.fields[f].pred);
            wc[f] = ew.expr(sp.fields[f].contrib);
        }
        o << "extern \"C\" __global__ __launch_bounds__(256) void rf_seg_fold_wave(RfSegArgs L) {\n";
        o << "    const int lane = threadIdx.x & 63;\n";
        o << "    const int64_t nw = (int64_t)gridDim.x * 4;\n";
        o << "    for (int64_t run = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); run < L.runs;\n"
             "         run += nw) {\n";
        o << "        const int64_t rr = run;\n";
        if (sp.ranged)
            o << "        const int64_t key = L.key0 + rr;\n        (void)key;\n";
        else
            o << "        const " << cty(sp.key_type) << " key = ((const " << cty(sp.key_type)
              << "*)L.keys)[rr];\n        (void)key;\n";
        o << init.str();
        o << "        {\n";
        if (sp.ranged) {
            o << bnd.str();
            o << "            int64_t lo = (int64_t)" << lo_e << ";\n";
            o << "            int64_t hi = (int64_t)" << hi_e << ";\n";
            o << "            lo = lo < 0 ? 0 : lo;\n";
            o << "            hi = hi < L.rows ? hi : L.rows;\n";
        } else {
            o << "            const int64_t lo = run > 0 ? (int64_t)L.run_ends[run - 1] : 0;\n";
            o << "            const int64_t hi0 = (int64_t)L.run_ends[run];\n";
            o << "            const int64_t hi = hi0 < L.rows ? hi0 : L.rows;\n";
        }
        for (size_t f = 0; f < sp.fields.size(); f++) {
            const std::string T = cty(sp.fields[f].type);
            o << "            " << T << " s" << f << " = (" << T << ")" << iv[f] << ";\n";
        }
        o << "            for (int64_t c = lo; c < hi; c += 64) {\n";
        o << "                const int64_t row0 = c + lane;\n";
        o << "                const int64_t row = row0 < hi ? row0 : lo;  // lanes past the end read row lo\n";
        o << wb.str();
        for (size_t f = 0; f < sp.fields.size(); f++) {
            if (!sp.fields[f].updated) continue;
            const std::string T = cty(sp.fields[f].type), F = std::to_string(f);
            o << "                const bool wp" << F << " = (bool)(" << wp[f] << ");\n";
            o << "                const " << T << " wv" << F << " = (" << T << ")(" << wc[f] << ");\n";
        }
        o << "                const int cnt = (int)(hi - c < 64 ? hi - c : 64);\n";
        o << "                for (int k = 0; k < cnt; k++) {\n";
        for (size_t f = 0; f < sp.fields.size(); f++) {
            if (!sp.fields[f].updated) continue;
            const std::string T = cty(sp.fields[f].type), S = "s" + std::to_string(f);
            const std::string F = std::to_string(f);
            o << "                    {\n"
              << "                        const bool pk = rf_readlane(wp" << F << ", k);\n"
              << "                        const " << T << " vk = rf_readlane(wv" << F << ", k);\n"
              << "                        " << S << " = pk ? "
              << combine_expr(sp.fields[f].op, sp.fields[f].type, S, "vk") << " : " << S << ";\n"
              << "                    }\n";
        }
        o << "                }\n";
        o << "            }\n";
        o << ob.str();
        o << "            if (lane == 0) {\n";
        for (size_t j = 0; j < ov.size(); j++) {
            const std::string T = cty(sp.out_types[j]);
            o << "                ((" << T << "*)L.out[" << j << "])[run] = (" << T << ")" << ov[j]
              << ";\n";
        }
        o << "            }\n        }\n    }\n}\n";
    }
    ks.code = o.str();
    return ks;
}

// The bits an LDS slot of field gf starts from: a float sum from -0.0 (the
// exact additive identity), an integer sum from 0, min/max from the fold
// identity.
std::string gj_slot_identity(const GField& gf) {
    if (gf.op == AggOp::ADD) {
        if (gf.type.bt == BT::FLOAT64) return "0x8000000000000000ull";
        if (gf.type.bt == BT::FLOAT32) return "0x80000000ull";
        return "0ull";
    }
    const std::string id = fold_identity(gf.op, gf.type);
    return gf.type.is_unsigned() ? "(unsigned long long)" + id
                                 : "(unsigned long long)(long long)" + id;
}

// Payload columns the group fold reads.
static std::set<int> gj_fold_cols(const GroupJoinPlan& gj) {
    std::set<int> used;
    for (auto& gf : gj.fields) {
        if (!gf.updated) continue;
        if (!gf.const_contrib) used.insert(gf.pay);
        if (gf.pred_pay >= 0) used.insert(gf.pred_pay);
    }
    return used;
}

// Rows per block of the wave fold's LDS buffer: 256, or 128 with more than
// four payload columns (four waves' buffers within 32 KiB).
int gj_fold_wave_rows(const GroupJoinPlan& gj) { return gj_fold_cols(gj).size() <= 4 ? 256 : 128; }

// Bytes of one wave's buffer (each column's block, 8-byte aligned).
static size_t gj_fold_wave_bytes(const GroupJoinPlan& gj) {
    size_t off = 0;
    const int R = gj_fold_wave_rows(gj);
    for (int j : gj_fold_cols(gj)) off += ((size_t)R * gj.pay_types[j].width() + 7) / 8 * 8;
    return std::max<size_t>(off, 8);
}

// Waves per block of rf_gj_lds_done: its LDS totals (8 bytes per slot,
// 4 per wave, and group of 16) within 32 KiB.
int gj_lds_done_waves(int nslot) { return nslot <= 4 ? 16 : (nslot <= 8 ? 8 : 4); }

bool gj_agg_ok(const GroupJoinPlan& gj) {
    for (const GField& gf : gj.fields) {
        if (!gf.updated) continue;
        const bool integer = gf.type.is_int() || gf.type.is_idx();
        if (gf.op == AggOp::ADD && (integer || gf.type.is_float())) continue;
        if ((gf.op == AggOp::MIN || gf.op == AggOp::MAX) && integer) continue;
        return false;
    }
    return gj.fields.size() <= 16;
}

KernelSource gen_group_fold(const GroupJoinPlan& gj, const std::vector<std::vector<bool>>& nullable) {
    KernelSource ks;
    ks.entry = "rf_gj_fold";
    ks.tune.block = 256;
    // Build-table columns are read by index at the group's build row.
    Gen g(gj.build_plan, nullable, ks);
    if (gj.out_values.size() > 16) throw Unsupported("more than 16 group join output columns");
    for (auto& t : gj.out_types) ks.out_types.push_back(t);
    // The fold of one group, in two forms with the same arithmetic in the
    // same order (the group's joined rows ascending, from the init):
    //  * rf_gj_fold_one: one thread folds its group;
    //  * rf_gj_fold_wave: the 64 lanes of a wave load the group's rows 64 at
    //    a time (coalesced) and every lane folds the chunk in row order from
    //    lane broadcasts -- for large groups, where a lone thread would wait
    //    on one load after another.
    auto fold_fn = [&](bool wave) {
        std::ostringstream o;
        o << "__device__ __forceinline__ void " << (wave ? "rf_gj_fold_wave" : "rf_gj_fold_one")
          << "(const RfFoldArgs& L, const int64_t grp" << (wave ? ", unsigned char* wbuf" : "")
          << ") {\n    {\n";
        if (wave)
            o << "        const int64_t i0 = rf_scalar64(L.seg_lo[grp]), i1 = rf_scalar64(L.seg_hi[grp]);\n";
        else
            o << "        const int64_t i0 = L.seg_lo[grp], i1 = L.seg_hi[grp];\n";
        std::ostringstream init;
        Emitter ei(g, Mode::FINAL, init, "        ");
        std::vector<std::string> iv;
        for (auto& f : gj.fields) iv.push_back(ei.expr(f.init));
        o << init.str();
        for (size_t f = 0; f < gj.fields.size(); f++) {
            const std::string T = cty(gj.fields[f].type);
            o << "        " << T << " s" << f << " = (" << T << ")" << iv[f] << ";\n";
        }
        std::ostringstream cst;
        Emitter ec(g, Mode::FINAL, cst, "        ");
        ec.ntmp = ei.ntmp;  // one scope: temporaries stay distinct
        std::vector<std::string> cv(gj.fields.size());
        for (size_t f = 0; f < gj.fields.size(); f++)
            if (gj.fields[f].updated && gj.fields[f].const_contrib)
                cv[f] = ec.expr(gj.fields[f].const_contrib);
        o << cst.str();
        // Per-field update from the row's payload values `val(j)`.
        auto updates = [&](const std::string& ind, const std::function<std::string(int)>& val) {
            for (size_t f = 0; f < gj.fields.size(); f++) {
                const GField& gf = gj.fields[f];
                if (!gf.updated) continue;
                const std::string T = cty(gf.type), S = "s" + std::to_string(f);
                std::string v = gf.const_contrib ? "(" + T + ")" + cv[f] : val(gf.pay);
                std::string upd = combine_expr(gf.op, gf.type, S, v);
                if (gf.pred_pay >= 0)
                    o << ind << S << " = " << val(gf.pred_pay) << " ? " << upd << " : " << S << ";\n";
                else
                    o << ind << S << " = " << upd << ";\n";
            }
        };
        if (!wave) {
            o << "#pragma unroll 4\n        for (int64_t i = i0; i < i1; i++) {\n";
            updates("            ", [&](int j) {
                return "((const " + cty(gj.pay_types[j]) + "*)L.pay[" + std::to_string(j) + "])[i]";
            });
            o << "        }\n";
        } else {
            std::set<int> used;
            for (auto& gf : gj.fields) {
                if (!gf.updated) continue;
                if (!gf.const_contrib) used.insert(gf.pay);
                if (gf.pred_pay >= 0) used.insert(gf.pred_pay);
            }
            // Blocks of R rows through the wave's LDS buffer: the lanes load a
            // block coalesced (64 rows per load), store it to LDS and issue
            // the next block's loads; then every lane folds the block in row
            // order, each row an LDS broadcast (same address in every lane).
            // The group's bounds are wave-uniform, so the row loop is a
            // scalar loop. (Rows broadcast from lanes with readlane, in a
            // loop whose bound the compiler took for a per-lane one: 513 us
            // for 512 groups of 11K rows.) Same additions, same order.
            const int R = gj_fold_wave_rows(gj);
            const std::string RS = std::to_string(R), DS = std::to_string(R / 64);
            o << "        const int lane = (int)(threadIdx.x & 63);\n";
            size_t off = 0;
            for (int j : used) {
                const std::string T = cty(gj.pay_types[j]), J = std::to_string(j);
                o << "        " << T << "* b" << J << " = (" << T << "*)(wbuf + " << off << ");\n";
                o << "        " << T << " r" << J << "[" << DS << "];\n";
                off += ((size_t)R * gj.pay_types[j].width() + 7) / 8 * 8;
            }
            auto loadR = [&](const std::string& at, const std::string& ind) {
                o << ind << "#pragma unroll\n" << ind << "for (int d = 0; d < " << DS << "; d++) {\n"
                  << ind << "    const int64_t i = " << at << " + d * 64 + lane;\n";
                for (int j : used) {
                    const std::string T = cty(gj.pay_types[j]), J = std::to_string(j);
                    o << ind << "    r" << J << "[d] = i < i1 ? ((const " << T << "*)L.pay[" << J
                      << "])[i] : (" << T << ")0;\n";
                }
                o << ind << "}\n";
            };
            loadR("i0", "        ");
            o << "        for (int64_t c0 = i0; c0 < i1; c0 += " << RS << ") {\n";
            o << "            #pragma unroll\n            for (int d = 0; d < " << DS << "; d++) {\n";
            for (int j : used) o << "                b" << j << "[d * 64 + lane] = r" << j << "[d];\n";
            o << "            }\n";
            loadR("c0 + " + RS, "            ");
            o << "            __builtin_amdgcn_wave_barrier();\n";
            o << "            const int n = (int)(i1 - c0 < " << RS << " ? i1 - c0 : " << RS << ");\n";
            o << "            #pragma unroll 16\n            for (int k = 0; k < n; k++) {\n";
            for (int j : used) {
                const std::string T = cty(gj.pay_types[j]), J = std::to_string(j);
                o << "                const " << T << " v" << J << " = b" << J << "[k];\n";
            }
            updates("                ", [&](int j) { return "v" + std::to_string(j); });
            o << "            }\n            __builtin_amdgcn_wave_barrier();\n        }\n";
        }
        o << "        const int64_t key = L.keys[grp];\n        const int64_t brow = L.brow[grp];\n"
             "        (void)key; (void)brow;\n";
        std::ostringstream body;
        Emitter em(g, Mode::FINAL, body, "        ");
        em.ntmp = ec.ntmp;
        std::vector<std::string> ov;
        for (auto& v : gj.out_values) ov.push_back(em.expr(v));
        o << body.str();
        if (wave) o << "        if (lane != 0) return;\n";
        for (size_t j = 0; j < ov.size(); j++) {
            const std::string T = cty(gj.out_types[j]);
            o << "        ((" << T << "*)L.out[" << j << "])[grp] = (" << T << ")" << ov[j] << ";\n";
        }
        o << "    }\n}\n";
        return o.str();
    };
    // rf_gj_agg (the default mode, gj_agg_ok fields): a block per partition of
    // 2^shift groups folds the partition's joined rows into LDS accumulators
    // with LDS atomics (one 8-byte slot per group and updated field; float
    // sums from -0.0, the exact additive identity), then evaluates each
    // group's state as init (+) accumulated total and its outputs.
    auto agg_fn = [&]() {
        std::ostringstream o;
        o << "extern \"C\" __global__ __launch_bounds__(256) void rf_gj_agg(RfFoldArgs L, RfAggArgs G) {\n"
             "    extern __shared__ unsigned long long rf_acc[];\n"
             "    const int64_t ngroups = G.groups_dev ? *G.groups_dev : L.groups;\n"
             "    const int64_t gpp = (int64_t)1 << G.shift;\n"
             "    const int64_t g0 = (int64_t)blockIdx.x << G.shift;\n"
             "    if (g0 >= ngroups) return;  // no group (and so no joined row) in this partition\n"
             "    const int64_t ng = ngroups - g0 < gpp ? ngroups - g0 : gpp;\n";
        std::vector<int> slot(gj.fields.size(), -1);
        int nslot = 0;
        for (size_t f = 0; f < gj.fields.size(); f++)
            if (gj.fields[f].updated) slot[f] = nslot++;
        auto ident_bits = [&](const GField& gf) { return gj_slot_identity(gf); };
        o << "    for (int64_t i = threadIdx.x; i < gpp; i += 256) {\n";
        for (size_t f = 0; f < gj.fields.size(); f++)
            if (slot[f] >= 0)
                o << "        rf_acc[" << slot[f] << " * gpp + i] = " << ident_bits(gj.fields[f])
                  << ";\n";
        o << "    }\n    __syncthreads();\n";
        // Constant contributions (COUNT's 1) are kernel parameters.
        std::ostringstream cst;
        Emitter ec(g, Mode::FINAL, cst, "    ");
        std::vector<std::string> cv(gj.fields.size());
        for (size_t f = 0; f < gj.fields.size(); f++)
            if (gj.fields[f].updated && gj.fields[f].const_contrib)
                cv[f] = ec.expr(gj.fields[f].const_contrib);
        o << cst.str();
        o << "    const int64_t r0 = G.pstart[blockIdx.x * G.stride];\n"
             "    const int64_t r1 = G.pstart[(blockIdx.x + 1) * G.stride];\n"
             "    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {\n"
             "        const int64_t g = (int64_t)G.prank[i] - g0;\n";
        auto val = [&](int j) {
            return "((const " + cty(gj.pay_types[j]) + "*)L.pay[" + std::to_string(j) + "])[i]";
        };
        for (size_t f = 0; f < gj.fields.size(); f++) {
            const GField& gf = gj.fields[f];
            if (slot[f] < 0) continue;
            const std::string T = cty(gf.type);
            const std::string v = gf.const_contrib ? "(" + T + ")" + cv[f] : "(" + T + ")" + val(gf.pay);
            const std::string a = "rf_acc[" + std::to_string(slot[f]) + " * gpp + g]";
            std::string upd;
            if (gf.op == AggOp::ADD && gf.type.bt == BT::FLOAT64)
                upd = "atomicAdd((double*)&" + a + ", (double)" + v + ")";
            else if (gf.op == AggOp::ADD && gf.type.bt == BT::FLOAT32)
                upd = "atomicAdd((float*)&" + a + ", (float)" + v + ")";
            else if (gf.op == AggOp::ADD)
                upd = "atomicAdd(&" + a + ", (unsigned long long)(long long)" + v + ")";
            else {
                const char* fn = gf.op == AggOp::MIN ? "atomicMin" : "atomicMax";
                upd = gf.type.is_unsigned()
                          ? std::string(fn) + "(&" + a + ", (unsigned long long)" + v + ")"
                          : std::string(fn) + "((long long*)&" + a + ", (long long)" + v + ")";
            }
            if (gf.pred_pay >= 0)
                o << "        if (" << val(gf.pred_pay) << ") " << upd << ";\n";
            else
                o << "        " << upd << ";\n";
        }
        o << "    }\n    __syncthreads();\n"
             "    for (int64_t gl = threadIdx.x; gl < ng; gl += 256) {\n"
             "        const int64_t grp = g0 + gl;\n";
        std::ostringstream init;
        Emitter ei(g, Mode::FINAL, init, "        ");
        ei.ntmp = ec.ntmp;
        std::vector<std::string> iv;
        for (auto& f : gj.fields) iv.push_back(ei.expr(f.init));
        o << init.str();
        for (size_t f = 0; f < gj.fields.size(); f++) {
            const GField& gf = gj.fields[f];
            const std::string T = cty(gf.type), S = "s" + std::to_string(f);
            o << "        " << T << " " << S << " = (" << T << ")" << iv[f] << ";\n";
            if (slot[f] < 0) continue;
            const std::string a = "rf_acc[" + std::to_string(slot[f]) + " * gpp + gl]";
            std::string tot;
            if (gf.type.bt == BT::FLOAT64) tot = "((const double*)rf_acc)[" + std::to_string(slot[f]) + " * gpp + gl]";
            else if (gf.type.bt == BT::FLOAT32) tot = "*(const float*)&" + a;
            else tot = "(" + T + ")(long long)" + a;
            o << "        " << S << " = " << combine_expr(gf.op, gf.type, S, tot) << ";\n";
        }
        o << "        const int64_t key = L.keys[grp];\n        const int64_t brow = L.brow[grp];\n"
             "        (void)key; (void)brow;\n";
        std::ostringstream body;
        Emitter em(g, Mode::FINAL, body, "        ");
        em.ntmp = ei.ntmp;
        std::vector<std::string> ov;
        for (auto& v : gj.out_values) ov.push_back(em.expr(v));
        o << body.str();
        for (size_t j = 0; j < ov.size(); j++) {
            const std::string T = cty(gj.out_types[j]);
            o << "        ((" << T << "*)L.out[" << j << "])[grp] = (" << T << ")" << ov[j] << ";\n";
        }
        o << "    }\n}\n";
        return o.str();
    };
    // rf_gj_lds_done (after a probe scan that pre-aggregated in LDS): a
    // block of W waves per 16 groups (gj_lds_done_waves); lane l of wave w
    // takes group 16*block + (l & 15) and folds the slots of blocks
    // 4w + (l >> 4), then every 4W-th (16 consecutive groups: two cache
    // lines per load), the 4W partial totals of a group combine in LDS, and
    // lanes 0-15 of wave 0 evaluate each group's state as init (+) total and
    // its outputs. (64 groups per block left most CUs idle: 20 us for the
    // date dimension's 2556 groups against 512 probe blocks.)
    auto lds_done_fn = [&]() {
        std::ostringstream o;
        std::vector<int> slot(gj.fields.size(), -1);
        int nslot = 0;
        for (size_t f = 0; f < gj.fields.size(); f++)
            if (gj.fields[f].updated) slot[f] = nslot++;
        const int W = gj_lds_done_waves(nslot);
        const int P = 4 * W;  // partial totals per group
        o << "extern \"C\" __global__ __launch_bounds__(" << W * 64
          << ") void rf_gj_lds_done(RfFoldArgs L, RfAggArgs G) {\n"
             "    __shared__ unsigned long long s_tot[" << std::max(nslot, 1) << "][" << P << "][16];\n"
             "    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);\n"
             "    const int gl = lane & 15, part = wv * 4 + (lane >> 4);\n"
             "    const int64_t ngroups = G.groups_dev ? *G.groups_dev : L.groups;\n"
             "    const int64_t grp = (int64_t)blockIdx.x * 16 + gl;\n"
             "    const bool live = grp < ngroups;\n";
        // Typed value of slot bits, and the bits of a typed value.
        auto typed = [&](const GField& gf, const std::string& bits) -> std::string {
            const std::string T = cty(gf.type);
            if (gf.type.bt == BT::FLOAT64) return "__longlong_as_double((long long)(" + bits + "))";
            if (gf.type.bt == BT::FLOAT32) return "__int_as_float((int)(unsigned)(" + bits + "))";
            return "(" + T + ")(long long)(" + bits + ")";
        };
        auto as_bits = [&](const GField& gf, const std::string& v) -> std::string {
            if (gf.type.bt == BT::FLOAT64) return "(unsigned long long)__double_as_longlong(" + v + ")";
            if (gf.type.bt == BT::FLOAT32) return "(unsigned long long)(unsigned)__float_as_int(" + v + ")";
            return "(unsigned long long)(long long)(" + v + ")";
        };
        for (size_t f = 0; f < gj.fields.size(); f++) {
            const GField& gf = gj.fields[f];
            if (slot[f] < 0) continue;
            const std::string T = cty(gf.type), S = std::to_string(slot[f]);
            o << "    {\n        " << T << " t = " << typed(gf, gj_slot_identity(gf)) << ";\n"
              << "        if (live)\n"
              << "            for (int64_t b = part; b < G.nblk; b += " << P << ")\n"
              << "                t = " << combine_expr(gf.op, gf.type, "t",
                                                     typed(gf, "G.bpart[(b * " + std::to_string(nslot) +
                                                                   " + " + S + ") * G.gpp + grp]"))
              << ";\n"
              << "        s_tot[" << S << "][part][gl] = " << as_bits(gf, "t") << ";\n    }\n";
        }
        o << "    __syncthreads();\n"
             "    if (wv != 0 || lane >= 16 || !live) return;\n"
             "    {\n";
        std::ostringstream init;
        Emitter ei(g, Mode::FINAL, init, "        ");
        std::vector<std::string> iv;
        for (auto& f : gj.fields) iv.push_back(ei.expr(f.init));
        o << init.str();
        for (size_t f = 0; f < gj.fields.size(); f++) {
            const GField& gf = gj.fields[f];
            const std::string T = cty(gf.type), S = "s" + std::to_string(f);
            o << "        " << T << " " << S << " = (" << T << ")" << iv[f] << ";\n";
            if (slot[f] < 0) continue;
            const std::string SL = std::to_string(slot[f]);
            o << "        {\n            " << T << " t = " << typed(gf, gj_slot_identity(gf)) << ";\n"
              << "            for (int w = 0; w < " << P << "; w++) t = "
              << combine_expr(gf.op, gf.type, "t", typed(gf, "s_tot[" + SL + "][w][gl]")) << ";\n"
              << "            " << S << " = " << combine_expr(gf.op, gf.type, S, "t") << ";\n        }\n";
        }
        o << "        const int64_t key = L.keys[grp];\n        const int64_t brow = L.brow[grp];\n"
             "        (void)key; (void)brow;\n";
        std::ostringstream body;
        Emitter em(g, Mode::FINAL, body, "        ");
        em.ntmp = ei.ntmp;
        std::vector<std::string> ov;
        for (auto& v : gj.out_values) ov.push_back(em.expr(v));
        o << body.str();
        for (size_t j = 0; j < ov.size(); j++) {
            const std::string T = cty(gj.out_types[j]);
            o << "        ((" << T << "*)L.out[" << j << "])[grp] = (" << T << ")" << ov[j] << ";\n";
        }
        o << "    }\n}\n";
        return o.str();
    };
    std::ostringstream o;
    o << "// Generated group fold and outputs for group join '" << gj.sym_name << "'.\n";
    o << fold_fn(false) << fold_fn(true);
    if (gj_agg_ok(gj)) o << agg_fn() << lds_done_fn();
    // A run's groups (a thread, or a wave, per group) and several runs' groups
    // in one launch (the run by binary search over cumulative group counts).
    const char* find_run =
        "        int64_t lo = 0, hi = B.n - 1;\n"
        "        while (lo < hi) {\n"
        "            const int64_t mid = (lo + hi + 1) >> 1;\n"
        "            if (B.gstart[mid] <= g) lo = mid; else hi = mid - 1;\n"
        "        }\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_gj_fold(RfFoldArgs L) {\n"
         "    for (int64_t grp = blockIdx.x * (int64_t)256 + threadIdx.x; grp < L.groups;\n"
         "         grp += (int64_t)gridDim.x * 256)\n"
         "        rf_gj_fold_one(L, grp);\n"
         "}\n";
    // The wave folds' LDS buffers: one per wave of the block.
    const std::string wlds = "    __shared__ unsigned long long rf_fold_lds[4][" +
                             std::to_string(gj_fold_wave_bytes(gj) / 8) +
                             "];\n    unsigned char* wbuf = (unsigned char*)rf_fold_lds[threadIdx.x >> 6];\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_gj_fold_w(RfFoldArgs L) {\n"
      << wlds
      << "    for (int64_t grp = rf_scalar64(blockIdx.x * (int64_t)4 + (threadIdx.x >> 6)); grp < L.groups;\n"
         "         grp += (int64_t)gridDim.x * 4)\n"
         "        rf_gj_fold_wave(L, grp, wbuf);\n"
         "}\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_gj_fold_b(RfFoldBatch B) {\n"
         "    const int64_t total = B.gstart[B.n];\n"
         "    for (int64_t g = blockIdx.x * (int64_t)256 + threadIdx.x; g < total;\n"
         "         g += (int64_t)gridDim.x * 256) {\n"
      << find_run
      << "        rf_gj_fold_one(*(const RfFoldArgs*)(B.list + lo), g - B.gstart[lo]);\n"
         "    }\n"
         "}\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void rf_gj_fold_wb(RfFoldBatch B) {\n"
      << wlds
      << "    const int64_t total = B.gstart[B.n];\n"
         "    for (int64_t g = rf_scalar64(blockIdx.x * (int64_t)4 + (threadIdx.x >> 6)); g < total;\n"
         "         g += (int64_t)gridDim.x * 4) {\n"
      << find_run
      << "        rf_gj_fold_wave(*(const RfFoldArgs*)(B.list + rf_scalar64(lo)), g - B.gstart[lo], wbuf);\n"
         "    }\n"
         "}\n";
    ks.code = o.str();
    return ks;
}

}  // namespace rf
