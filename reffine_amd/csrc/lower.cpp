// GPU lowering: Func -> Plan.  See lower.hpp for the contract.
#include "lower.hpp"

#include <algorithm>
#include <functional>
#include <map>
#include <set>
#include <sstream>

namespace rf {

std::string leaf_name(const N& n) {
    if (n && n->kind == K::Sym && !n->name.empty() && n->name[0] == '$')
        return n->name;
    return "";
}

bool is_leaf(const N& n, const char* prefix) {
    auto s = leaf_name(n);
    return !s.empty() && s.rfind(prefix, 0) == 0;
}

static N leaf(const std::string& name, const Type& t) { return mk_sym(name, t); }

static bool same(const N& a, const N& b) {
    if (a == b) return true;
    if (!a || !b) return false;
    if (a->kind == K::Sym && b->kind == K::Sym) return a->name == b->name && a->type == b->type;
    return false;
}

// Structural equality of two expressions.
static bool same_tree(const N& a, const N& b) {
    if (a == b) return true;
    if (!a || !b) return false;
    if (a->kind != b->kind || !(a->type == b->type) || a->op != b->op || a->name != b->name ||
        a->val != b->val || a->idx != b->idx || a->args.size() != b->args.size() ||
        a->iters.size() != b->iters.size())
        return false;
    for (size_t i = 0; i < a->args.size(); i++)
        if (!same_tree(a->args[i], b->args[i])) return false;
    for (size_t i = 0; i < a->iters.size(); i++)
        if (!same_tree(a->iters[i], b->iters[i])) return false;
    return true;
}

static bool mentions(const N& e, const std::function<bool(const N&)>& p) {
    if (!e) return false;
    if (p(e)) return true;
    for (auto& a : e->args)
        if (mentions(a, p)) return true;
    for (auto& a : e->iters)
        if (mentions(a, p)) return true;
    return false;
}

static bool mentions_state(const N& e) {
    return mentions(e, [](const N& n) { return is_leaf(n, "$s."); });
}

static N mk_and(N a, N b) {
    if (!a) return b;
    if (!b) return a;
    if (a->kind == K::Const && a->val != 0) return b;
    if (b->kind == K::Const && b->val != 0) return a;
    return mk_nary(MathOp::AND, {a, b});
}

static N mk_or(N a, N b) { return mk_nary(MathOp::OR, {a, b}); }

static N true_c() { return mk_const(types::BOOL_T, 1); }

// ---------------------------------------------------------------- context

struct Lowerer {
    N func;
    bool vectorize;
    std::map<std::string, N> tbl;        // table symbol name -> expr
    std::map<std::string, int> input_ix; // input symbol name -> index
    Plan plan;

    // Row binding: what the scan's row means.
    N row_val_sym;      // Reduce value symbol bound to a SubVector row
    N iter_sym;         // Op iterator symbol (VecSpace / range form)
    N merge_iter[2];    // MERGE: each side's _in iterator expression (t or t + c)
    int driver = -1;
    int driver_dim = 1;
    // An In(t, T) over this input drives the scan when present (the runs
    // table of a run stage); otherwise the first In does.
    int prefer_driver = -1;
    std::map<std::string, int> probe_of;  // probe key text -> probe index

    int input_index(const N& sym) {
        auto it = input_ix.find(sym->name);
        return it == input_ix.end() ? -1 : it->second;
    }

    N col_leaf(int table, int col) {
        const Type& vt = plan.inputs[table]->type;
        if (col < 0 || col >= (int)vt.dtypes.size())
            throw IRError("column index out of range");
        if (col < (int)vt.enc.size() && vt.enc[col] == Enc::RUNEND) {
            // ReadData of a run-end child reads its values child at the row
            // index (src/pass/readwritepass.cpp:16-29): the value of run
            // `row`. Rows at or past the number of runs read 0 here (the
            // reference reads them unchecked).
            return mk_readdata(leaf("$tab." + std::to_string(table), vt),
                               leaf("$row", types::IDX_T), col);
        }
        return leaf("$col." + std::to_string(table) + "." + std::to_string(col),
                    vt.dtypes[col]);
    }
    N valid_leaf(int table, int col) {
        return leaf("$valid." + std::to_string(table) + "." + std::to_string(col),
                    types::BOOL_T);
    }

    // Row of a table as the struct reffine's Element/SubVector produce:
    // columns dim.. of a dim-1 table (type.h:132-142 valty).
    N row_struct(int table, int first_col) {
        const Type& vt = plan.inputs[table]->type;
        std::vector<N> cols;
        for (int c = first_col; c < (int)vt.dtypes.size(); c++) cols.push_back(col_leaf(table, c));
        return mk_new(cols);
    }

    int probe_for(int table, N key) {
        std::string k = std::to_string(table) + "|" + to_text(key);
        auto it = probe_of.find(k);
        if (it != probe_of.end()) return it->second;
        const Type& vt = plan.inputs[table]->type;
        if (vt.dim != 1) throw Unsupported("probe into a multi-dimensional vector");
        if (vt.dtypes[0] != types::I64)
            throw Unsupported("hash probe needs an int64 iterator column "
                              "(build_index, include/reffine/arrow/table.h:107-138)");
        if (key->type != types::I64) throw IRError("probe key must be i64");
        int j = (int)plan.probes.size();
        plan.probes.push_back({table, key});
        probe_of[k] = j;
        return j;
    }

    // Probed row: columns of table T at the row holding `key` in column 0.
    N probe_row(int table, N key) {
        int j = probe_for(table, key);
        const Type& vt = plan.inputs[table]->type;
        std::vector<N> cols;
        for (int c = vt.dim; c < (int)vt.dtypes.size(); c++)
            cols.push_back(leaf("$pcol." + std::to_string(j) + "." + std::to_string(c),
                                vt.dtypes[c]));
        return mk_new(cols);
    }

    // ------------------------------------------------------ the rewriter

    std::map<Node*, N> memo;
    std::map<std::string, N> subst;  // symbol name -> replacement
    std::map<std::string, N> reduce_subst;  // Reduce text -> its result leaves

    // A lowerer over the same function, with no row binding yet.
    Lowerer sibling() const {
        Lowerer L;
        L.func = func;
        L.vectorize = vectorize;
        L.tbl = tbl;
        L.input_ix = input_ix;
        L.plan.func_name = plan.func_name;
        L.plan.inputs = plan.inputs;
        L.plan.tbl = plan.tbl;
        L.plan.vectorize = vectorize;
        return L;
    }

    N rw(const N& e) {
        auto m = memo.find(e.get());
        if (m != memo.end()) return m->second;
        N r = rw_(e);
        memo[e.get()] = r;
        return r;
    }

    N rw_(const N& e) {
        switch (e->kind) {
            case K::Sym: {
                if (!leaf_name(e).empty()) return e;
                auto s = subst.find(e->name);
                if (s != subst.end()) return s->second;
                int ix = input_index(e);
                if (ix >= 0) {
                    if (e->type.is_vector())
                        throw Unsupported("vector input '" + e->name +
                                          "' used as a row value");
                    return leaf("$arg." + std::to_string(ix), e->type);
                }
                auto t = tbl.find(e->name);
                if (t != tbl.end()) return rw(t->second);
                throw IRError("undefined symbol " + e->name);
            }
            case K::Const:
                return e;
            case K::Cast:
                return mk_cast(e->type, rw(e->args[0]));
            case K::Select: {
                N c = rw(e->args[0]), t = rw(e->args[1]), f = rw(e->args[2]);
                if (t->kind == K::New && f->kind == K::New) {
                    std::vector<N> v;
                    for (size_t i = 0; i < t->args.size(); i++)
                        v.push_back(simplify_select(c, t->args[i], f->args[i]));
                    return mk_new(v);
                }
                return simplify_select(c, t, f);
            }
            case K::Get: {
                N v = rw(e->args[0]);
                return get_of(v, e->idx);
            }
            case K::New: {
                std::vector<N> v;
                for (auto& a : e->args) v.push_back(rw(a));
                return mk_new(v);
            }
            case K::Nary: {
                std::vector<N> v;
                for (auto& a : e->args) v.push_back(rw(a));
                return mk_nary(e->op, v);
            }
            case K::Element:
                return rw_element(e);
            case K::In:
                return rw_in(e);
            case K::Length: {
                N vec = resolve_vec(e->args[0]);
                int t = input_index(vec);
                if (t < 0) throw Unsupported("length of a computed vector");
                return leaf("$len." + std::to_string(t) + "." + std::to_string(e->idx),
                            types::IDX_T);
            }
            case K::InitVal:
                // The initialised symbols are lowered where they are used.
                return rw(e->args[0]);
            case K::BuildIdx:
                throw Unsupported("an index over a computed vector outside the symbol table");
            case K::Locate: {
                // vector_locate (include/reffine/vinstr/vinstr.h:137-146): a
                // probe of the table's index; -1 when the key is absent.
                N vec = resolve_vec(e->args[0]);
                int t = input_index(vec);
                if (t < 0) throw Unsupported("Locate in a computed vector");
                int j = probe_for(t, rw(e->args[1]));
                return leaf("$prow." + std::to_string(j), types::IDX_T);
            }
            case K::ReadData: {
                // ReadData(T, i, c): column c (storage index) at row i
                // (src/pass/readwritepass.cpp:16-24); here i must be a row
                // found by Locate in the same table, or the scanned row.
                N vec = resolve_vec(e->args[0]);
                int t = input_index(vec);
                if (t < 0) throw Unsupported("ReadData of a computed vector");
                const Type& vt = plan.inputs[t]->type;
                const int c = (int)e->idx;
                if (c < (int)vt.enc.size() && vt.enc[c] == Enc::RUNEND)
                    throw Unsupported("ReadData of a run-end encoded column");
                N ix = rw(e->args[1]);
                const std::string ln = leaf_name(ix);
                if (ln.rfind("$prow.", 0) == 0) {
                    int j = std::stoi(ln.substr(6));
                    if (plan.probes.at(j).table != t)
                        throw Unsupported("ReadData at a row located in another table");
                    return leaf("$pcol." + std::to_string(j) + "." + std::to_string(c),
                                vt.dtypes[c]);
                }
                if (ln == "$row" && t == driver && plan.source == Plan::TABLE_ROWS)
                    return col_leaf(t, c);
                // Any other row index: an indexed read of the column (rows
                // outside the table read 0; the reference reads them
                // unchecked).
                return mk_readdata(leaf("$tab." + std::to_string(t), vt), ix, c);
            }
            case K::Reduce: {
                auto r = reduce_subst.find(to_text(e));
                if (r != reduce_subst.end()) return r->second;
                throw Unsupported("reduction inside a row expression: " + to_text(e));
            }
            default:
                throw Unsupported("node kind inside a row expression: " + to_text(e));
        }
    }

    N simplify_select(N c, N t, N f) {
        if (t->type.is_struct()) {
            // Select of structs that are not literal New nodes.
            std::vector<N> v;
            for (size_t i = 0; i < t->type.dtypes.size(); i++)
                v.push_back(simplify_select(c, get_of(t, i), get_of(f, i)));
            return mk_new(v);
        }
        if (same(t, f)) return t;
        return mk_select(c, t, f);
    }

    N get_of(N v, int64_t k) {
        if (v->kind == K::New) {
            return v->args.at(k);
        }
        if (v->kind == K::Select) return simplify_select(v->args[0], get_of(v->args[1], k),
                                                         get_of(v->args[2], k));
        if (v->type.is_primitive() && k == 0) return v;  // canonpass.cpp:22-29
        return mk_get(v, k);
    }

    N resolve_vec(const N& e) {
        if (e->kind == K::Sym) {
            if (input_index(e) >= 0) return e;
            auto t = tbl.find(e->name);
            if (t != tbl.end()) return resolve_vec(t->second);
        }
        return e;
    }

    N rw_element(const N& e) {
        N vec = resolve_vec(e->args[0]);
        int t = input_index(vec);
        if (t < 0) throw Unsupported("Element of a computed vector");
        const N& it = e->args[1];
        if (plan.source == Plan::MERGE && (t == plan.merge_a || t == plan.merge_b)) {
            // A side's element at its iteration expression (t, or t + c for
            // a shifted side) reads that side's current row
            // (JointSpace::_vec_iter_idxs, iter_space.cpp:252-262; LoopGen
            // Element, loopgen.cpp:12-20: the index map is keyed by that
            // expression), whatever that row's key. At another expression
            // the reference's map lookup fails.
            const int side = t == plan.merge_a ? 0 : 1;
            N w = resolve_bool_sym(it);
            if (!same_tree(w, merge_iter[side]))
                throw Unsupported("an element of a merged vector read at an expression other "
                                  "than its _in's iterator expression");
            const std::string S = side == 0 ? "0" : "1";
            const Type& vt = plan.inputs[t]->type;
            std::vector<N> cols;
            for (size_t c = 1; c < vt.dtypes.size(); c++)
                cols.push_back(leaf("$mcol." + S + "." + std::to_string(c), vt.dtypes[c]));
            return mk_new(cols);
        }
        if (iter_sym && it->kind == K::Sym && it->name == iter_sym->name &&
            t == driver) {
            // Iterator-bound element: direct row (loopgen.cpp:16-17).
            if (plan.inputs[t]->type.dim != 1)
                throw Unsupported("element of a multi-dimensional vector");
            return row_struct(t, 1);
        }
        // Otherwise a hash probe (loopgen.cpp:18-19, vector_locate).
        return probe_row(t, rw(it));
    }

    N rw_in(const N& e) {
        N vec = resolve_vec(e->args[1]);
        int t = input_index(vec);
        if (t < 0) throw Unsupported("In over a computed vector");
        if (t == driver && iter_sym && e->args[0]->kind == K::Sym &&
            e->args[0]->name == iter_sym->name)
            return true_c();  // the scan itself
        int j = probe_for(t, rw(e->args[0]));
        return leaf("$found." + std::to_string(j), types::BOOL_T);
    }

    // -------------------------------------------- state-field decomposition

    struct Dec {
        bool upd = false;
        AggOp op = AggOp::ADD;
        N pred, contrib;
    };

    Dec decompose(const N& E, const std::string& fname) {
        Dec d;
        if (E->kind == K::Sym && E->name == fname) return d;  // no update
        if (E->kind == K::Select) {
            const N& c = E->args[0];
            if (mentions_state(c))
                throw Unsupported("reduction predicate depends on the state");
            Dec t = decompose(E->args[1], fname), f = decompose(E->args[2], fname);
            if (!t.upd && !f.upd) return d;
            if (!t.upd) {
                f.pred = mk_and(mk_nary(MathOp::NOT, {c}), f.pred);
                return f;
            }
            if (!f.upd) {
                t.pred = mk_and(c, t.pred);
                return t;
            }
            if (t.op != f.op) throw Unsupported("select between different folds");
            d.upd = true;
            d.op = t.op;
            N tp = t.pred ? t.pred : true_c(), fp = f.pred ? f.pred : true_c();
            d.pred = (t.pred || f.pred)
                         ? mk_select(c, tp, fp)
                         : nullptr;
            d.contrib = mk_select(c, t.contrib, f.contrib);
            return d;
        }
        if (E->kind == K::Nary && E->args.size() == 2) {
            const N& a = E->args[0];
            const N& b = E->args[1];
            bool a_s = a->kind == K::Sym && a->name == fname;
            bool b_s = b->kind == K::Sym && b->name == fname;
            AggOp op;
            bool comm = true;  // the state may be either operand
            bool neg = false;  // s - x folds as s + (-x)
            switch (E->op) {
                case MathOp::ADD: op = AggOp::ADD; break;
                case MathOp::SUB: op = AggOp::ADD; comm = false; neg = true; break;
                case MathOp::MUL: op = AggOp::MUL; break;
                // std::max(x, s) / std::min(x, s) with the state second keep
                // x on a NaN or a tie of zeros: a float fold of that form is
                // not the state-first fold the kernels re-associate.
                case MathOp::MAX: op = AggOp::MAX; comm = !E->type.is_float(); break;
                case MathOp::MIN: op = AggOp::MIN; comm = !E->type.is_float(); break;
                case MathOp::AND: op = AggOp::AND; break;
                case MathOp::OR: op = AggOp::OR; break;
                default:
                    throw Unsupported("state field updated by " + mathop_name(E->op));
            }
            if (a_s && !mentions_state(b)) {
                d.upd = true;
                d.op = op;
                d.contrib = neg ? mk_nary(MathOp::NEG, {b}) : b;
                return d;
            }
            if (b_s && comm && !mentions_state(a)) {
                d.upd = true;
                d.op = op;
                d.contrib = a;
                return d;
            }
        }
        throw Unsupported("reduction state field is not an associative fold: " +
                          to_text(E));
    }

    // Flattened state leaves for a state type; returns New-tree of leaves.
    N state_tree(const Type& t, std::vector<int>& path, std::vector<Field>& fields) {
        if (t.is_struct()) {
            std::vector<N> v;
            for (size_t i = 0; i < t.dtypes.size(); i++) {
                path.push_back((int)i);
                v.push_back(state_tree(t.dtypes[i], path, fields));
                path.pop_back();
            }
            return mk_new(v);
        }
        if (!t.is_primitive()) throw Unsupported("non-primitive reduction state field");
        Field f;
        f.type = t;
        f.path = path;
        std::string name = "$s." + std::to_string(fields.size());
        fields.push_back(f);
        return leaf(name, t);
    }

    static void flatten(const N& e, std::vector<N>& out) {
        if (e->kind == K::New) {
            for (auto& a : e->args) flatten(a, out);
        } else {
            if (e->type.is_struct()) throw Unsupported("unflattenable struct value");
            out.push_back(e);
        }
    }

    // ------------------------------------------------ scan source setup

    // Binds the scan to a SubVector (loopgen.cpp:221-242) or an Op
    // (loopgen.cpp:44-111). Returns the row value struct and the row
    // predicate.
    struct Source {
        N row_value;  // struct value delivered to acc / Op outputs
        N pred;       // row predicate (validity, Op predicate)
    };

    Source bind_subvec(const N& sv) {
        N vec = resolve_vec(sv->args[0]);
        int t = input_index(vec);
        if (t < 0) throw Unsupported("SubVector of a computed vector");
        set_driver(t);
        plan.source = Plan::TABLE_ROWS;
        plan.start = sv->args[1];
        plan.end = sv->args[2];
        const Type& vt = plan.inputs[t]->type;
        if (vt.dtypes.size() < 2) throw Unsupported("SubVector of a one-column vector");
        Source s;
        // Outputs are ReadData of columns 1.. and the body condition is the
        // validity of column 1 (loopgen.cpp:231-238).
        s.row_value = row_struct(t, 1);
        s.pred = valid_leaf(t, 1);
        return s;
    }

    void set_driver(int t) {
        if (driver >= 0 && driver != t)
            throw Unsupported("components scan different tables");
        driver = t;
        plan.driver = t;
    }

    // Splits an Op predicate into conjuncts.
    static void conjuncts(const N& e, std::vector<N>& out) {
        if (e->kind == K::Nary && e->op == MathOp::AND) {
            conjuncts(e->args[0], out);
            conjuncts(e->args[1], out);
        } else {
            out.push_back(e);
        }
    }

    N resolve_bool_sym(const N& e) {
        if (e->kind == K::Sym && leaf_name(e).empty() && input_index(e) < 0) {
            auto t = tbl.find(e->name);
            if (t != tbl.end()) return resolve_bool_sym(t->second);
        }
        return e;
    }

    // Bounds lb <= it < ub of iterator `it` among the conjuncts (removed from
    // `cj`); false if it is not bounded on both sides.
    bool take_bounds(const N& it, std::vector<N>& cj, N& lb, N& ub) {
        std::vector<N> rest;
        lb = ub = nullptr;
        auto one = mk_const(it->type, 1);
        for (auto& c0 : cj) {
            N c = resolve_bool_sym(c0);
            if (c->kind == K::Nary && c->args.size() == 2 &&
                (c->op == MathOp::GTE || c->op == MathOp::GT || c->op == MathOp::LT ||
                 c->op == MathOp::LTE)) {
                bool l_it = c->args[0]->kind == K::Sym && c->args[0]->name == it->name;
                bool r_it = c->args[1]->kind == K::Sym && c->args[1]->name == it->name;
                if (l_it && !r_it) {
                    N b = c->args[1];
                    if (c->op == MathOp::GTE && !lb) { lb = b; continue; }
                    if (c->op == MathOp::GT && !lb) { lb = mk_nary(MathOp::ADD, {b, one}); continue; }
                    if (c->op == MathOp::LT && !ub) { ub = b; continue; }
                    if (c->op == MathOp::LTE && !ub) { ub = mk_nary(MathOp::ADD, {b, one}); continue; }
                } else if (r_it && !l_it) {
                    N b = c->args[0];
                    if (c->op == MathOp::LTE && !lb) { lb = b; continue; }
                    if (c->op == MathOp::LT && !lb) { lb = mk_nary(MathOp::ADD, {b, one}); continue; }
                    if (c->op == MathOp::GT && !ub) { ub = b; continue; }
                    if (c->op == MathOp::GTE && !ub) { ub = mk_nary(MathOp::ADD, {b, one}); continue; }
                }
            }
            rest.push_back(c);
        }
        if (!lb || !ub) return false;
        cj = rest;
        return true;
    }

    static bool mentions_sym(const N& e, const std::string& name) {
        return mentions(e, [&](const N& n) { return n->kind == K::Sym && n->name == name; });
    }

    // Two iterators bounded by ranges (NestedSpace of two bounded spaces,
    // test/src/test_nested.cpp:9-25): one range scan over the product,
    // row r = (a - lb_a) * n_b + (b - lb_b), a outer and b inner as the
    // nested loops emit them.
    Source bind_nested(const N& op) {
        const N a = op->iters[0], b = op->iters[1];
        for (auto& it : op->iters)
            if (!it->type.is_int() && !it->type.is_idx())
                throw Unsupported("nested Op over a non-integer iterator");
        std::vector<N> cj;
        conjuncts(resolve_bool_sym(op->args[0]), cj);
        for (auto& c0 : cj)
            if (resolve_bool_sym(c0)->kind == K::In)
                throw Unsupported("two-iterator Op over a table");
        N lba, uba, lbb, ubb;
        if (!take_bounds(a, cj, lba, uba) || !take_bounds(b, cj, lbb, ubb))
            throw Unsupported("Op iterators are not bounded on both sides");
        for (auto& e : {lba, uba, lbb, ubb})
            if (mentions_sym(e, a->name) || mentions_sym(e, b->name))
                throw Unsupported("an iterator bound depends on another iterator");
        plan.source = Plan::RANGE;
        const N zero = mk_const(types::IDX_T, 0);
        N na = mk_nary(MathOp::MAX, {mk_nary(MathOp::SUB, {mk_cast(types::IDX_T, uba),
                                                           mk_cast(types::IDX_T, lba)}), zero});
        N nb = mk_nary(MathOp::MAX, {mk_nary(MathOp::SUB, {mk_cast(types::IDX_T, ubb),
                                                           mk_cast(types::IDX_T, lbb)}), zero});
        plan.start = zero;
        plan.end = mk_nary(MathOp::MUL, {na, nb});
        // Row-side expressions: bounds are host values (constants or scalar
        // inputs), rewritten to leaves.
        const N row = leaf("$row", types::IDX_T);
        const N nbr = rw(nb);
        const N ia = mk_nary(MathOp::ADD, {mk_cast(types::IDX_T, rw(lba)),
                                           mk_nary(MathOp::DIV, {row, nbr})});
        const N ib = mk_nary(MathOp::ADD, {mk_cast(types::IDX_T, rw(lbb)),
                                           mk_nary(MathOp::MOD, {row, nbr})});
        subst[a->name] = mk_cast(a->type, ia);
        subst[b->name] = mk_cast(b->type, ib);
        iter_sym = nullptr;
        Source s;
        N pred;
        for (auto& c : cj) pred = mk_and(pred, rw(c));
        s.pred = pred ? pred : true_c();
        std::vector<N> vals;
        for (auto& it : op->iters) vals.push_back(rw(it));
        for (size_t i = 1; i < op->args.size(); i++) vals.push_back(rw(op->args[i]));
        s.row_value = mk_new(vals);
        return s;
    }

    // Finds the conjunct _in(t, A) | _in(t, B) (t the Op's iterator, A and B
    // two input tables keyed by int64): sets the plan's merge source, binds
    // t to the merge's current key and removes the conjunct. Other unions of
    // iteration spaces are refused by name.
    bool bind_merge(const N& op, std::vector<N>& cj) {
        int found = -1;
        for (size_t k = 0; k < cj.size(); k++) {
            N c = resolve_bool_sym(cj[k]);
            if (!(c->kind == K::Nary && c->op == MathOp::OR)) continue;
            bool has_in = false;
            std::function<void(const N&)> f = [&](const N& x) {
                N y = resolve_bool_sym(x);
                if (y->kind == K::In) has_in = true;
                if (y->kind == K::Nary && y->op == MathOp::OR)
                    for (auto& a : y->args) f(a);
            };
            f(c);
            if (!has_in) continue;  // a plain boolean filter
            // A side is _in(t, X) or _in(t + c, X) with c a constant: the
            // Reffine pass makes the latter a ShiftedSpace (reffinepass.cpp:
            // 67-71) whose iterator at X's row i is X.key[i] + c
            // (iter_space.cpp:120-140).
            int64_t offs[2] = {0, 0};
            N its[2];
            auto side = [&](const N& x, int s) -> int {
                N y = resolve_bool_sym(x);
                if (y->kind != K::In) return -1;
                N e = resolve_bool_sym(y->args[0]);
                auto is_iter = [&](const N& v) {
                    return v->kind == K::Sym && v->name == iter_sym->name;
                };
                if (e->kind == K::Nary && e->op == MathOp::ADD && e->args.size() == 2 &&
                    is_iter(e->args[0])) {
                    N cst = resolve_bool_sym(e->args[1]);
                    if (cst->kind != K::Const || cst->type.bt != BT::INT64) return -1;
                    offs[s] = (int64_t)cst->val;  // constants are held as doubles
                } else if (!is_iter(e)) {
                    return -1;
                }
                its[s] = e;
                return input_index(resolve_vec(y->args[1]));
            };
            const int a = side(c->args[0], 0), b = side(c->args[1], 1);
            if (a < 0 || b < 0)
                throw Unsupported("a union of iteration spaces other than _in(t, A) | _in(t, B) "
                                  "over two input tables (t the Op's iterator, or t + c for a "
                                  "constant c)");
            if (a == b) throw Unsupported("a union of a vector with itself");
            if (found >= 0) throw Unsupported("an Op over more than one union");
            for (int t : {a, b}) {
                const Type& vt = plan.inputs[t]->type;
                if (vt.dim != 1) throw Unsupported("a union merge over a multi-dimensional vector");
                if (vt.dtypes[0].bt != BT::INT64 || iter_sym->type.bt != BT::INT64)
                    throw Unsupported("a union merge needs int64 keys and iterator");
                for (size_t col = 0; col < vt.enc.size(); col++)
                    if (vt.enc[col] == Enc::RUNEND)
                        throw Unsupported("a run-end encoded column in a union merge");
            }
            found = (int)k;
            plan.merge_a = a;
            plan.merge_b = b;
            plan.merge_off[0] = offs[0];
            plan.merge_off[1] = offs[1];
            merge_iter[0] = its[0];
            merge_iter[1] = its[1];
        }
        if (found < 0) return false;
        if (op->iters.size() != 1) throw Unsupported("a union merge with more than one iterator");
        cj.erase(cj.begin() + found);
        plan.source = Plan::MERGE;
        plan.start = mk_const(types::IDX_T, 0);
        plan.end = mk_nary(MathOp::ADD,
                           {leaf("$len." + std::to_string(plan.merge_a) + ".0", types::IDX_T),
                            leaf("$len." + std::to_string(plan.merge_b) + ".0", types::IDX_T)});
        subst[iter_sym->name] = leaf("$mkey", iter_sym->type);
        return true;
    }

    Source bind_op(const N& op) {
        if (op->iters.size() == 2) return bind_nested(op);
        if (op->iters.size() != 1)
            throw Unsupported("Op with more than two iterators");
        iter_sym = op->iters[0];
        std::vector<N> cj;
        conjuncts(resolve_bool_sym(op->args[0]), cj);
        // The first In(t, T) over an input table drives the scan (VecSpace,
        // iter_space.cpp:55-91); further In's are hash semi-joins
        // (InterSpace, iter_space.cpp:318-364).
        if (bind_merge(op, cj)) {
            // _in(t, A) | _in(t, B): UnionSpace (iter_space.cpp:216-316), the
            // outer merge of two vectors ascending by key. The loop's body
            // condition is UnionSpace::_iter_cond (either side's current row
            // is in range with a valid key); further conjuncts filter
            // (FilteredSpace over the union).
            N pred = mk_or(leaf("$mvalid.0.0", types::BOOL_T), leaf("$mvalid.1.0", types::BOOL_T));
            for (auto& c : cj) pred = mk_and(pred, rw(c));
            Source s;
            s.pred = pred;
            std::vector<N> vals;
            for (auto& it : op->iters) vals.push_back(rw(it));
            for (size_t i = 1; i < op->args.size(); i++) vals.push_back(rw(op->args[i]));
            s.row_value = mk_new(vals);
            return s;
        }
        int drv = -1;
        for (auto& c0 : cj) {
            N c = resolve_bool_sym(c0);
            if (c->kind == K::In && c->args[0]->kind == K::Sym &&
                c->args[0]->name == iter_sym->name) {
                N vec = resolve_vec(c->args[1]);
                int t = input_index(vec);
                if (t >= 0 && (drv < 0 || t == prefer_driver)) drv = t;
                if (drv >= 0 && (prefer_driver < 0 || drv == prefer_driver)) break;
            }
        }
        Source s;
        N pred;
        if (drv >= 0) {
            set_driver(drv);
            plan.source = Plan::TABLE_ROWS;
            const Type& vt = plan.inputs[drv]->type;
            if (vt.dim != 1) throw Unsupported("Op over a multi-dimensional vector");
            if (vt.dtypes[0] != iter_sym->type) throw IRError("iterator type mismatch");
            plan.start = mk_const(types::IDX_T, 0);
            plan.end = leaf("$len." + std::to_string(drv) + ".0", types::IDX_T);
            subst[iter_sym->name] = col_leaf(drv, 0);
            // VecSpace::_iter_cond: validity of column 0 (iter_space.cpp:60-66)
            pred = valid_leaf(drv, 0);
        } else {
            // Range form: iterator bounded by lb <= t < ub (gen_table_op,
            // test/src/test_utils.cpp:32-46).
            plan.source = Plan::RANGE;
            if (!iter_sym->type.is_int() && !iter_sym->type.is_idx())
                throw Unsupported("range Op over a non-integer iterator");
            N lb, ub;
            std::vector<N> rest;
            for (auto& c0 : cj) {
                N c = resolve_bool_sym(c0);
                if (c->kind == K::Nary && c->args.size() == 2 &&
                    (c->op == MathOp::GTE || c->op == MathOp::GT ||
                     c->op == MathOp::LT || c->op == MathOp::LTE)) {
                    bool l_it = c->args[0]->kind == K::Sym && c->args[0]->name == iter_sym->name;
                    bool r_it = c->args[1]->kind == K::Sym && c->args[1]->name == iter_sym->name;
                    auto one = mk_const(iter_sym->type, 1);
                    if (l_it && !r_it) {
                        N b = c->args[1];
                        if (c->op == MathOp::GTE && !lb) { lb = b; continue; }
                        if (c->op == MathOp::GT && !lb) { lb = mk_nary(MathOp::ADD, {b, one}); continue; }
                        if (c->op == MathOp::LT && !ub) { ub = b; continue; }
                        if (c->op == MathOp::LTE && !ub) { ub = mk_nary(MathOp::ADD, {b, one}); continue; }
                    } else if (r_it && !l_it) {
                        N b = c->args[0];
                        if (c->op == MathOp::LTE && !lb) { lb = b; continue; }
                        if (c->op == MathOp::LT && !lb) { lb = mk_nary(MathOp::ADD, {b, one}); continue; }
                        if (c->op == MathOp::GT && !ub) { ub = b; continue; }
                        if (c->op == MathOp::GTE && !ub) { ub = mk_nary(MathOp::ADD, {b, one}); continue; }
                    }
                }
                rest.push_back(c);
            }
            if (!lb || !ub) throw Unsupported("Op iterator is not bounded on both sides");
            plan.start = mk_cast(types::IDX_T, lb);
            plan.end = mk_cast(types::IDX_T, ub);
            subst[iter_sym->name] = mk_cast(iter_sym->type, leaf("$row", types::IDX_T));
            cj = rest;
        }
        for (auto& c : cj) pred = mk_and(pred, rw(c));
        s.pred = pred ? pred : true_c();
        std::vector<N> vals;
        for (auto& it : op->iters) vals.push_back(rw(it));
        for (size_t i = 1; i < op->args.size(); i++) vals.push_back(rw(op->args[i]));
        s.row_value = mk_new(vals);
        return s;
    }

    // ------------------------------------------------------ components

    ReduceComp lower_reduce(const N& red, const std::string& sym_name) {
        const N& vec = red->args[0];
        const N& init = red->args[1];
        const N& ssym = red->args[2];
        const N& vsym = red->args[3];
        const N& acc = red->args[4];
        if (vec->type.dim != 1)  // loopgen.cpp:209
            throw Unsupported("reduction over a multi-dimensional vector");
        N rv = resolve_vec(vec);
        Source src;
        if (rv->kind == K::SubVector) {
            src = bind_subvec(rv);
        } else if (rv->kind == K::Op) {
            src = bind_op(rv);
        } else {
            throw Unsupported("reduction over " + to_text(rv));
        }
        ReduceComp rc;
        rc.state_type = red->type;
        rc.sym_name = sym_name;
        std::vector<int> path;
        N stree = state_tree(red->type, path, rc.fields);
        subst[ssym->name] = stree;
        subst[vsym->name] = src.row_value;
        memo.clear();
        N accr = rw(acc);
        std::vector<N> flat;
        flatten(accr, flat);
        if (flat.size() != rc.fields.size()) throw IRError("accumulator shape");
        std::vector<N> inits;
        flatten(fold_host(init), inits);
        for (size_t f = 0; f < rc.fields.size(); f++) rc.fields[f].init = inits.at(f);
        try {
            for (size_t f = 0; f < rc.fields.size(); f++) {
                Dec d = decompose(flat[f], "$s." + std::to_string(f));
                Field& fl = rc.fields[f];
                fl.updated = d.upd;
                if (d.upd) {
                    fl.op = d.op;
                    // vectorize: state = body_cond ? acc : state; otherwise
                    // if (body_cond) state = acc -- the same fold
                    // (loopgen.cpp:253-262).
                    fl.pred = mk_and(src.pred, d.pred);
                    if (!fl.pred) fl.pred = true_c();
                    fl.contrib = d.contrib;
                    if (mentions_state(fl.contrib) || mentions_state(fl.pred))
                        throw Unsupported("state field depends on another field");
                }
            }
        } catch (const Unsupported&) {
            // Not re-associable: the reference's own loop, row by row
            // (loopgen.cpp:253-262), run by one wave (rf_qs).
            rc.sequential = true;
            rc.seq_pred = src.pred;
            rc.seq_next = flat;
            for (auto& fl : rc.fields) {
                fl.updated = true;
                fl.pred = nullptr;
                fl.contrib = nullptr;
            }
        }
        return rc;
    }

    // Structural inlining for host-side expressions (init values, bounds).
    N fold_host(const N& e) {
        memo.clear();
        return rw(e);
    }

    OpComp lower_op(const N& op, const std::string& sym_name) {
        Source src = bind_op(op);
        OpComp oc;
        oc.sym_name = sym_name;
        oc.pred = src.pred;
        oc.values = src.row_value->args;
        oc.n_iters = (int)op->iters.size();
        for (auto& v : oc.values) {
            if (!v->type.is_primitive()) throw Unsupported("non-primitive Op output");
            oc.types.push_back(v->type);
        }
        oc.vec_type = op->type;
        return oc;
    }

    // ------------------------------------------------------- group join

    // Reductions reachable from an expression through the symbol table.
    void find_reduces(const N& e, std::vector<N>& out, std::set<std::string>& seen) {
        if (!e) return;
        if (e->kind == K::Sym) {
            if (!leaf_name(e).empty() || input_index(e) >= 0) return;
            auto t = tbl.find(e->name);
            if (t != tbl.end()) find_reduces(t->second, out, seen);
            return;
        }
        if (e->kind == K::Reduce) {
            if (seen.insert(to_text(e)).second) out.push_back(e);
            return;
        }
        for (auto& a : e->args) find_reduces(a, out, seen);
    }

    bool is_group_join(const N& op) {
        std::vector<N> reds;
        std::set<std::string> seen;
        for (size_t i = 1; i < op->args.size(); i++) find_reduces(op->args[i], reds, seen);
        return !reds.empty();
    }

    static bool is_sym_named(const N& e, const std::string& name) {
        return e->kind == K::Sym && e->name == name;
    }

    void lower_group_join(const N& op, const std::string& name) {
        if (op->iters.size() != 1) throw Unsupported("group join with more than one iterator");
        std::vector<N> reds;
        std::set<std::string> seen;
        for (size_t i = 1; i < op->args.size(); i++) find_reduces(op->args[i], reds, seen);
        if (reds.size() != 1)
            throw Unsupported("a group join needs exactly one inner reduction");
        const N red = reds[0];
        const N t = op->iters[0];
        auto gj = std::make_shared<GroupJoinPlan>();

        // Build side: the outer Op's iteration space (VecSpace over B plus
        // its predicate), selecting (key, build row).
        Lowerer Lb = sibling();
        Source sb = Lb.bind_op(mk_op(op->iters, op->args[0], {}));
        if (Lb.plan.source != Plan::TABLE_ROWS)
            throw Unsupported("group join whose outer Op is not over a table");
        const int B = Lb.driver;
        if (plan.inputs[B]->type.dtypes[0] != types::I64)
            throw Unsupported("group key must be an int64 column 0");
        if (!Lb.plan.probes.empty())
            throw Unsupported("probes in the group join's build predicate");
        OpComp bo;
        bo.sym_name = name + ".build";
        bo.pred = sb.pred;
        bo.values = {Lb.col_leaf(B, 0), mk_cast(types::I64, leaf("$row", types::IDX_T))};
        bo.types = {types::I64, types::I64};
        bo.vec_type = Type(BT::VECTOR, {types::I64, types::I64}, 1);
        Lb.plan.ops = {bo};
        // The qualifying keys' range, folded in the same scan (MIN, MAX of
        // the key under the selection's predicate): the rank map is sized
        // from it.
        ReduceComp kr;
        kr.sym_name = name + ".keyrange";
        kr.state_type = Type(BT::STRUCT, {types::I64, types::I64});
        for (AggOp op : {AggOp::MIN, AggOp::MAX}) {
            Field f;
            f.type = types::I64;
            f.op = op;
            f.updated = true;
            f.pred = sb.pred;
            f.contrib = Lb.col_leaf(B, 0);
            f.init = mk_const(types::I64, 0);
            f.path = {(int)kr.fields.size()};
            kr.fields.push_back(f);
        }
        Lb.plan.reduces = {kr};
        Lb.plan.func_name = func->name + "_build";
        Lb.plan.output_type = bo.vec_type;
        gj->build = B;

        // Probe side: the inner reduction's Op without the key equality.
        N rv = resolve_vec(red->args[0]);
        if (rv->kind != K::Op || rv->iters.size() != 1)
            throw Unsupported("inner reduction of a group join must be over a one-iterator Op");
        std::vector<N> cj;
        conjuncts(resolve_bool_sym(rv->args[0]), cj);
        N key_expr, rest;
        for (auto& c0 : cj) {
            N c = resolve_bool_sym(c0);
            if (!key_expr && c->kind == K::Nary && c->op == MathOp::EQ && c->args.size() == 2) {
                const bool l = is_sym_named(c->args[0], t->name);
                const bool r = is_sym_named(c->args[1], t->name);
                if (l != r) {
                    key_expr = l ? c->args[1] : c->args[0];
                    continue;
                }
            }
            rest = rest ? mk_nary(MathOp::AND, {rest, c}) : c;
        }
        if (!key_expr)
            throw Unsupported("inner reduction has no equality with the outer iterator");
        std::vector<N> outs(rv->args.begin() + 1, rv->args.end());
        Lowerer Lp = sibling();
        Source sp;
        N key;
        try {
            sp = Lp.bind_op(mk_op(rv->iters, rest ? rest : true_c(), outs));
            key = Lp.rw(key_expr);
        } catch (const IRError& e) {
            throw Unsupported(std::string("inner reduction depends on the outer row beyond the "
                                          "key equality: ") + e.what());
        }
        if (Lp.plan.source != Plan::TABLE_ROWS)
            throw Unsupported("group join whose inner Op is not over a table");
        if (key->type != types::I64) key = mk_cast(types::I64, key);
        const int gjix = (int)Lp.plan.probes.size();
        if (gjix >= 4) throw Unsupported("more than 4 probed tables");
        Probe gp;
        gp.table = B;
        gp.key = key;
        gp.guard = sp.pred;
        gp.external = true;
        Lp.plan.probes.push_back(gp);
        const N rank = leaf("$prow." + std::to_string(gjix), types::I64);

        // Fields of the inner reduction.
        std::vector<Field> fields;
        std::vector<int> path;
        N stree = Lp.state_tree(red->type, path, fields);
        Lp.subst[red->args[2]->name] = stree;
        Lp.subst[red->args[3]->name] = sp.row_value;
        Lp.memo.clear();
        N accr;
        try {
            accr = Lp.rw(red->args[4]);
        } catch (const IRError& e) {
            throw Unsupported(std::string("group join accumulator depends on the outer row: ") +
                              e.what());
        }
        std::vector<N> flat;
        flatten(accr, flat);
        if (flat.size() != fields.size()) throw IRError("accumulator shape");
        std::vector<N> inits;
        flatten(fold_host(red->args[1]), inits);
        OpComp po;
        po.sym_name = name + ".probe";
        po.pred = mk_and(sp.pred, mk_nary(MathOp::GTE, {rank, mk_const(types::I64, 0)}));
        po.values = {mk_cast(types::I32, rank)};
        po.types = {types::I32};
        for (size_t f = 0; f < fields.size(); f++) {
            Dec d = Lp.decompose(flat[f], "$s." + std::to_string(f));
            GField g;
            g.type = fields[f].type;
            g.path = fields[f].path;
            g.init = inits.at(f);
            g.updated = d.upd;
            if (d.upd) {
                g.op = d.op;
                if (mentions_state(d.contrib) || (d.pred && mentions_state(d.pred)))
                    throw Unsupported("state field depends on another field");
                if (d.contrib->kind == K::Const) {
                    g.const_contrib = d.contrib;
                } else {
                    g.pay = (int)po.values.size() - 1;
                    po.values.push_back(d.contrib);
                    po.types.push_back(g.type);
                }
                if (d.pred) {
                    g.pred_pay = (int)po.values.size() - 1;
                    po.values.push_back(d.pred);
                    po.types.push_back(types::BOOL_T);
                }
            }
            gj->fields.push_back(g);
        }
        if (po.values.size() > 9) throw Unsupported("more than 8 group join payload columns");
        for (size_t j = 1; j < po.types.size(); j++) gj->pay_types.push_back(po.types[j]);
        std::vector<Type> vts;
        for (auto& ty : po.types) vts.push_back(ty);
        po.vec_type = Type(BT::VECTOR, vts, 1);
        Lp.plan.ops = {po};
        Lp.plan.early_sel = true;  // joined rows are a large share of the probe rows
        Lp.plan.func_name = func->name + "_probe";
        Lp.plan.output_type = po.vec_type;
        gj->probe = Lp.driver;
        gj->state_type = red->type;

        // Finalize: the outer Op's outputs per group.
        Lowerer Lf = sibling();
        Lf.set_driver(B);
        Lf.iter_sym = t;
        Lf.subst[t->name] = leaf("$key", types::I64);
        std::vector<N> accl;
        for (size_t f = 0; f < fields.size(); f++)
            accl.push_back(leaf("$acc." + std::to_string(f), fields[f].type));
        size_t k = 0;
        std::function<N(const Type&)> nest = [&](const Type& ty) -> N {
            if (ty.is_struct()) {
                std::vector<N> v;
                for (auto& d : ty.dtypes) v.push_back(nest(d));
                return mk_new(v);
            }
            return accl.at(k++);
        };
        Lf.reduce_subst[to_text(red)] = nest(red->type);
        gj->out_values.push_back(leaf("$key", types::I64));
        gj->out_types.push_back(types::I64);
        for (size_t i = 1; i < op->args.size(); i++) {
            N v = Lf.rw(op->args[i]);
            if (!v->type.is_primitive()) throw Unsupported("non-primitive group join output");
            gj->out_values.push_back(v);
            gj->out_types.push_back(v->type);
        }
        Lf.plan.func_name = func->name + "_final";
        gj->build_plan = Lb.plan;
        gj->probe_plan = Lp.plan;
        gj->vec_type = op->type;
        gj->sym_name = name;
        plan.gj = gj;
        plan.driver = gj->probe;
        plan.source = Plan::TABLE_ROWS;
        plan.start = mk_const(types::IDX_T, 0);
        plan.end = leaf("$len." + std::to_string(gj->probe) + ".0", types::IDX_T);
    }

    // ------------------------------------------------------- run segments

    // T if `red` reduces T[t] (T a table input whose column 0 is run-end
    // encoded, t a symbol: the run of t's rows, loopgen.cpp:34-38); -1
    // otherwise.
    int run_table_of(const N& red, N* it = nullptr) {
        if (red->kind != K::Reduce) return -1;
        N v = resolve_vec(red->args[0]);
        if (v->kind != K::Element || v->args[1]->kind != K::Sym) return -1;
        int t = input_index(resolve_vec(v->args[0]));
        if (t < 0) return -1;
        const Type& vt = plan.inputs[t]->type;
        if (vt.dim < 2 || vt.enc.empty() || vt.enc[0] != Enc::RUNEND) return -1;
        if (it) *it = v->args[1];
        return t;
    }

    static void check_run_table(const Type& vt) {
        if (vt.dtypes.size() < 2) throw Unsupported("run-end table without value columns");
        for (size_t c = 1; c < vt.dtypes.size(); c++)
            if (c < vt.enc.size() && vt.enc[c] == Enc::RUNEND)
                throw Unsupported("run-end encoded value column");
    }

    // The reduction of an Op output whose vector is T[t], with t the Op's
    // iterator and the Op's space exactly In(t, T); null otherwise.
    N seg_reduce_of(const N& op) {
        if (op->iters.size() != 1) return nullptr;
        std::vector<N> reds;
        std::set<std::string> seen;
        for (size_t i = 1; i < op->args.size(); i++) find_reduces(op->args[i], reds, seen);
        if (reds.size() != 1) return nullptr;
        N it;
        const int T = run_table_of(reds[0], &it);
        if (T < 0 || it->name != op->iters[0]->name) return nullptr;
        std::vector<N> cj;
        conjuncts(resolve_bool_sym(op->args[0]), cj);
        for (auto& c0 : cj) {
            N c = resolve_bool_sym(c0);
            if (!(c->kind == K::In && is_sym_named(c->args[0], it->name) &&
                  input_index(resolve_vec(c->args[1])) == T))
                return nullptr;
        }
        return cj.empty() ? nullptr : reds[0];
    }

    // Appends the fields of one run reduction to `fields`: the fold over
    // the run's rows (columns 1.. of T) under the SubVector body condition,
    // the validity of column 1 (src/ir/op_to_loop.h:11-25 via
    // loopgen.cpp:34-38). The run's key may appear as $key (subst).
    void lower_run_fields(const N& red, int T, std::vector<Field>& fields) {
        const size_t off = fields.size();
        std::vector<int> path;
        N stree = state_tree(red->type, path, fields);
        subst[red->args[2]->name] = stree;
        subst[red->args[3]->name] = row_struct(T, 1);
        memo.clear();
        N accr;
        try {
            accr = rw(red->args[4]);
        } catch (const IRError& e) {
            throw Unsupported(std::string("run reduction depends on the run beyond its rows: ") +
                              e.what());
        }
        std::vector<N> flat;
        flatten(accr, flat);
        if (flat.size() != fields.size() - off) throw IRError("accumulator shape");
        std::vector<N> inits;
        flatten(fold_host(red->args[1]), inits);
        for (size_t f = 0; f < flat.size(); f++) {
            Dec d = decompose(flat[f], "$s." + std::to_string(off + f));
            Field& fl = fields[off + f];
            fl.init = inits.at(f);
            fl.updated = d.upd;
            if (d.upd) {
                fl.op = d.op;
                fl.pred = mk_and(valid_leaf(T, 1), d.pred);
                fl.contrib = d.contrib;
                if (mentions_state(fl.contrib) || mentions_state(fl.pred))
                    throw Unsupported("state field depends on another field");
            }
        }
        memo.clear();
    }

    // A reduction's state rebuilt from one leaf per flattened field.
    static N nest_fields(const Type& ty, const std::function<N(size_t)>& leaf_of, size_t& k) {
        if (ty.is_struct()) {
            std::vector<N> v;
            for (auto& d : ty.dtypes) v.push_back(nest_fields(d, leaf_of, k));
            return mk_new(v);
        }
        return leaf_of(k++);
    }

    void lower_seg_reduce(const N& op, const N& red, const std::string& name) {
        const N t = op->iters[0];
        const int T = run_table_of(red);
        const Type& vt = plan.inputs[T]->type;
        check_run_table(vt);
        auto sp = std::make_shared<SegReducePlan>();
        sp->table = T;
        sp->key_type = vt.dtypes[0];
        Lowerer Lr = sibling();
        Lr.set_driver(T);
        Lr.subst[t->name] = leaf("$key", t->type);
        Lr.lower_run_fields(red, T, sp->fields);
        // Outputs per run over the key and the fold's fields.
        size_t k = 0;
        Lr.reduce_subst[to_text(red)] = nest_fields(
            red->type,
            [&](size_t f) { return leaf("$acc." + std::to_string(f), sp->fields[f].type); }, k);
        sp->out_values.push_back(leaf("$key", t->type));
        sp->out_types.push_back(t->type);
        for (size_t i = 1; i < op->args.size(); i++) {
            N v = Lr.rw(op->args[i]);
            if (!v->type.is_primitive()) throw Unsupported("non-primitive run reduction output");
            sp->out_values.push_back(v);
            sp->out_types.push_back(v->type);
        }
        sp->sub = std::move(Lr.plan);
        sp->state_type = red->type;
        sp->vec_type = op->type;
        sp->sym_name = name;
        plan.seg = sp;
        plan.driver = T;
        plan.source = Plan::TABLE_ROWS;
        plan.start = mk_const(types::IDX_T, 0);
        plan.end = leaf("$len." + std::to_string(T) + ".0", types::IDX_T);
    }

    // A ranged fold: an Op over one range iterator (lb <= i < ub and nothing
    // else) whose outputs reduce _subvec(T, lo(i), hi(i)) of an input table
    // T, every reduction over the same rows -- misc.h:104-139's per-body
    // force sums over the pair table. One fold per i of rows [lo, hi), in
    // row order (the SubVector loop, loopgen.cpp:221-242).
    bool lower_range_fold(const N& op, const std::string& name) {
        if (op->iters.size() != 1) return false;
        std::vector<N> reds;
        std::set<std::string> seen;
        for (size_t i = 1; i < op->args.size(); i++) find_reduces(op->args[i], reds, seen);
        if (reds.empty()) return false;
        int T = -1;
        N lo, hi;
        for (auto& r : reds) {
            N v = resolve_vec(r->args[0]);
            if (v->kind != K::SubVector) return false;
            const int t = input_index(resolve_vec(v->args[0]));
            if (t < 0) return false;
            if (T >= 0 && (t != T || to_text(v->args[1]) != to_text(lo) ||
                           to_text(v->args[2]) != to_text(hi)))
                throw Unsupported("reductions over different ranges in one ranged fold");
            T = t;
            lo = v->args[1];
            hi = v->args[2];
        }
        for (auto& c0 : [&] { std::vector<N> cj; conjuncts(resolve_bool_sym(op->args[0]), cj); return cj; }())
            if (resolve_bool_sym(c0)->kind == K::In) return false;
        const N it = op->iters[0];
        const Type& vt = plan.inputs[T]->type;
        if (vt.dtypes.size() < 2) throw Unsupported("SubVector of a one-column vector");
        auto sp = std::make_shared<SegReducePlan>();
        sp->ranged = true;
        sp->table = T;
        sp->key_type = types::I64;
        Lowerer Lr = sibling();
        Lr.set_driver(T);
        Lr.subst[it->name] = mk_cast(it->type, leaf("$key", types::I64));
        std::vector<size_t> offs;
        for (auto& r : reds) {
            offs.push_back(sp->fields.size());
            Lr.lower_run_fields(r, T, sp->fields);
        }
        Lr.memo.clear();
        sp->lo = Lr.rw(lo);
        sp->hi = Lr.rw(hi);
        // Bounds and outputs: the Op bound as a range scan, its reductions
        // read as the fold's fields, its iterator ($row) as the fold's key.
        Lowerer Lb = sibling();
        for (size_t r = 0; r < reds.size(); r++) {
            size_t k = offs[r];
            Lb.reduce_subst[to_text(reds[r])] = nest_fields(
                reds[r]->type,
                [&](size_t f) { return leaf("$acc." + std::to_string(f), sp->fields[f].type); },
                k);
        }
        Source src = Lb.bind_op(op);
        if (Lb.plan.source != Plan::RANGE || !Lb.plan.probes.empty())
            throw Unsupported("ranged fold over a table scan or with probes");
        if (!(src.pred->kind == K::Const && src.pred->val != 0))
            throw Unsupported("ranged fold with a predicate beyond the iterator's bounds");
        sp->lb = Lb.plan.start;
        sp->ub = Lb.plan.end;
        std::function<N(const N&)> key_of = [&](const N& e) -> N {
            if (is_leaf(e, "$row")) return mk_cast(e->type, leaf("$key", types::I64));
            if (e->args.empty()) return e;
            auto c = std::make_shared<Node>(*e);
            for (auto& a : c->args) a = key_of(a);
            return c;
        };
        for (auto& v : src.row_value->args) {
            if (!v->type.is_primitive()) throw Unsupported("non-primitive ranged fold output");
            sp->out_values.push_back(key_of(v));
            sp->out_types.push_back(v->type.is_idx() ? types::I64 : v->type);
        }
        sp->sub = std::move(Lr.plan);
        sp->state_type = reds[0]->type;
        sp->vec_type = op->type;
        sp->sym_name = name;
        plan.seg = sp;
        plan.driver = T;
        plan.source = Plan::TABLE_ROWS;
        plan.start = mk_const(types::IDX_T, 0);
        plan.end = leaf("$len." + std::to_string(T) + ".1", types::IDX_T);
        return true;
    }

    // A run stage: reductions over T[t] (T's column 0 run-end encoded)
    // anywhere in the function, t bound by an Op whose space has In(t, T) --
    // the shape of TPC-H Q3/Q4/Q12/Q18 over lineitem (benchmark/include/
    // tpch.h:8-158, 528-723), where the reference merges T's runs with
    // another table (InterSpace, iter_space.cpp:318-364) and folds each
    // matching run. Here a first kernel folds every run of T into a runs
    // table (the run's key, then the fields of every such reduction; a
    // field may read other tables at the run's key), and the function is
    // lowered as usual with that table in T's place: In(t, T) scans it and
    // each reduction reads its fields at the scanned row.
    bool lower_run_stage(const N& out) {
        std::vector<N> reds;
        std::set<std::string> texts;
        int T = -1;
        N it;
        std::set<const Node*> seen;
        std::set<std::string> syms;
        std::function<void(const N&)> find = [&](const N& e) {
            if (!e || !seen.insert(e.get()).second) return;
            N i;
            const int t = run_table_of(e, &i);
            if (t >= 0) {
                if (T >= 0 && (t != T || i->name != it->name))
                    throw Unsupported("run reductions over different tables or iterators");
                T = t;
                it = i;
                if (texts.insert(to_text(e)).second) reds.push_back(e);
                return;
            }
            if (e->kind == K::Sym) {
                if (input_index(e) >= 0 || !leaf_name(e).empty()) return;
                if (!syms.insert(e->name).second) return;
                auto d = tbl.find(e->name);
                if (d != tbl.end()) find(d->second);
                return;
            }
            for (auto& a : e->args) find(a);
        };
        find(out);
        if (T < 0) return false;
        // T may appear only as In(t, T) and inside its run reductions.
        bool other_use = false;
        seen.clear();
        syms.clear();
        std::function<void(const N&)> uses = [&](const N& e) {
            if (!e || !seen.insert(e.get()).second) return;
            if (run_table_of(e) >= 0) return;
            if (e->kind == K::In && input_index(resolve_vec(e->args[1])) == T) {
                if (!is_sym_named(e->args[0], it->name)) other_use = true;
                return;
            }
            if (e->kind == K::Sym) {
                const int ix = input_index(e);
                if (ix == T) other_use = true;
                if (ix >= 0 || !leaf_name(e).empty()) return;
                if (!syms.insert(e->name).second) return;
                auto d = tbl.find(e->name);
                if (d != tbl.end()) uses(d->second);
                return;
            }
            for (auto& a : e->args) uses(a);
        };
        uses(out);
        if (other_use)
            throw Unsupported("run-end table used beyond In(t, T) and reductions over its runs");
        const Type& vt = plan.inputs[T]->type;
        check_run_table(vt);
        auto rs = std::make_shared<RunStagePlan>();
        rs->table = T;
        rs->input = plan.inputs[T];
        SegReducePlan& sp = rs->seg;
        sp.table = T;
        sp.key_type = vt.dtypes[0];
        Lowerer Lr = sibling();
        Lr.set_driver(T);
        Lr.subst[it->name] = leaf("$key", it->type);
        std::vector<size_t> offs;
        for (auto& r : reds) {
            offs.push_back(sp.fields.size());
            Lr.lower_run_fields(r, T, sp.fields);
        }
        if (sp.fields.size() + 1 > 16) throw Unsupported("more than 15 run reduction fields");
        sp.out_values.push_back(leaf("$key", it->type));
        sp.out_types.push_back(vt.dtypes[0]);
        for (size_t f = 0; f < sp.fields.size(); f++) {
            const Type& ft = sp.fields[f].type;
            if (ft.is_idx()) throw Unsupported("idx-typed field of a run reduction");
            sp.out_values.push_back(leaf("$acc." + std::to_string(f), ft));
            sp.out_types.push_back(ft);
        }
        sp.sub = std::move(Lr.plan);
        sp.sym_name = it->name;
        sp.vec_type = Type(BT::VECTOR, sp.out_types, 1,
                           std::vector<Enc>(sp.out_types.size(), Enc::FLAT));
        // The rest of the function scans the runs table in T's place.
        plan.inputs[T] = mk_sym(rs->input->name, sp.vec_type);
        prefer_driver = T;
        for (size_t r = 0; r < reds.size(); r++) {
            size_t k = offs[r];
            reduce_subst[to_text(reds[r])] = nest_fields(
                reds[r]->type, [&](size_t f) { return col_leaf(T, 1 + (int)f); }, k);
        }
        plan.runs = rs;
        return true;
    }

    Plan run() {
        plan.func_name = func->name;
        for (size_t i = 1; i < func->args.size(); i++) {
            plan.inputs.push_back(func->args[i]);
            input_ix[func->args[i]->name] = (int)i - 1;
        }
        for (auto& kv : func->tbl) tbl[kv.first->name] = kv.second;
        plan.tbl = func->tbl;
        plan.vectorize = vectorize;
        N out = func->args[0];
        plan.output_type = out->type;
        // Collect the components the output needs: Reduce / Op symbols.
        std::vector<std::pair<std::string, N>> comps;
        std::set<std::string> seen;
        std::function<void(const N&)> walk = [&](const N& e) {
            if (e->kind == K::Sym) {
                auto t = tbl.find(e->name);
                if (t == tbl.end()) return;
                N d = t->second;
                while (d->kind == K::InitVal) d = d->args[0];  // its inits are lowered where used
                if (d->kind == K::Reduce || d->kind == K::Op) {
                    if (seen.insert(e->name).second) comps.push_back({e->name, d});
                    return;
                }
                walk(t->second);
                return;
            }
            if (e->kind == K::Reduce || e->kind == K::Op)
                throw Unsupported("unnamed Reduce/Op in the output expression");
            for (auto& a : e->args) walk(a);
        };
        walk(out);
        if (comps.empty()) throw Unsupported("function output has no Reduce or Op");
        if (comps.size() == 1 && comps[0].second->kind == K::Op) {
            if (N red = seg_reduce_of(comps[0].second)) {
                if (!(out->kind == K::Sym && out->name == comps[0].first))
                    throw Unsupported("a run reduction must be the function's output");
                lower_seg_reduce(comps[0].second, red, comps[0].first);
                plan.output = out;
                return plan;
            }
        }
        if (comps.size() == 1 && comps[0].second->kind == K::Op &&
            out->kind == K::Sym && out->name == comps[0].first &&
            lower_range_fold(comps[0].second, comps[0].first)) {
            plan.output = out;
            return plan;
        }
        const bool staged = lower_run_stage(out);
        if (!staged && comps.size() == 1 && comps[0].second->kind == K::Op &&
            is_group_join(comps[0].second)) {
            if (!(out->kind == K::Sym && out->name == comps[0].first))
                throw Unsupported("a group join must be the function's output");
            lower_group_join(comps[0].second, comps[0].first);
            plan.output = out;
            return plan;
        }
        for (auto& [name, node] : comps) {
            if (node->kind == K::Reduce) {
                plan.reduces.push_back(lower_reduce(node, name));
            } else {
                if (!plan.ops.empty()) throw Unsupported("more than one Op output");
                plan.ops.push_back(lower_op(node, name));
            }
        }
        if (plan.source == Plan::MERGE &&
            (!plan.reduces.empty() || plan.ops.size() != 1 || staged))
            throw Unsupported("a union merge must be the function's only component (an Op "
                              "output)");
        if (staged && plan.driver != plan.runs->table)
            throw Unsupported("reductions over runs outside an Op over In(t, T)");
        for (auto& r : plan.reduces) {
            if (!r.sequential) continue;
            if (plan.reduces.size() != 1 || !plan.ops.empty() || staged)
                throw Unsupported("a fold that cannot be re-associated must be the function's "
                                  "only component");
            plan.sequential = true;
        }
        guard_probes();
        plan.output = out;
        return plan;
    }

    // A leaf of the scanned row that no probe feeds: its columns and their
    // validity, the row index, scalar arguments and lengths.
    bool probe_free(const N& e) const {
        const std::string drv = std::to_string(plan.driver) + ".";
        return !mentions(e, [&](const N& n) {
            if (n->kind != K::Sym) return false;
            const std::string ln = leaf_name(n);
            if (ln.empty()) return true;
            if (ln == "$row" || ln.rfind("$arg.", 0) == 0 || ln.rfind("$len.", 0) == 0)
                return false;
            if (ln.rfind("$col." + drv, 0) == 0 || ln.rfind("$valid." + drv, 0) == 0)
                return false;
            return true;
        });
    }

    // Every output of the scan (the Op's rows, each updated reduction field)
    // is taken only on rows whose predicate holds. The predicate's conjuncts
    // that read no probe and appear in every output's predicate guard the
    // probes: a row failing them is probed as "key absent" (-1), which no
    // output can see, and skips the index lookup's dependent loads.
    void guard_probes() {
        if (plan.probes.empty() || plan.sequential || plan.source != Plan::TABLE_ROWS) return;
        for (auto& pr : plan.probes)
            if (pr.table == plan.driver) return;
        std::vector<N> preds;
        for (auto& o : plan.ops) preds.push_back(o.pred);
        for (auto& r : plan.reduces)
            for (auto& f : r.fields)
                if (f.updated) preds.push_back(f.pred);
        if (preds.empty()) return;
        std::vector<std::vector<N>> cs(preds.size());
        for (size_t i = 0; i < preds.size(); i++) {
            if (!preds[i]) return;
            conjuncts(preds[i], cs[i]);
        }
        N guard;
        for (auto& c : cs[0]) {
            if (!probe_free(c)) continue;
            bool everywhere = true;
            for (size_t i = 1; i < cs.size() && everywhere; i++) {
                bool found = false;
                for (auto& d : cs[i]) found = found || same_tree(c, d);
                everywhere = found;
            }
            if (everywhere) guard = guard ? mk_nary(MathOp::AND, {guard, c}) : c;
        }
        if (!guard) return;
        for (auto& pr : plan.probes)
            if (!pr.external && !pr.guard) pr.guard = guard;
    }
};

static N strip_initval(N d) {
    while (d && d->kind == K::InitVal) d = d->args[0];
    return d;
}

N split_materialized(const N& func, std::vector<Prelude>& pre) {
    if (!func || func->kind != K::Func) throw IRError("expected a Func");
    std::map<std::string, N> tbl, tsym;
    for (auto& kv : func->tbl) {
        tbl[kv.first->name] = kv.second;
        tsym[kv.first->name] = kv.first;
    }
    std::set<std::string> inputs;
    for (size_t i = 1; i < func->args.size(); i++) inputs.insert(func->args[i]->name);
    std::vector<std::string> cands;
    std::set<std::string> cand_set;
    std::set<const Node*> seen;
    // A vector operand naming a computed vector (Op or BuildIndex).
    auto computed = [&](const N& v) -> bool {
        if (v->kind != K::Sym || inputs.count(v->name) || !tbl.count(v->name)) return false;
        N d = strip_initval(tbl[v->name]);
        if (d->kind != K::Op && d->kind != K::BuildIdx) return false;
        if (cand_set.insert(v->name).second) cands.push_back(v->name);
        return true;
    };
    std::function<void(const N&)> visit = [&](const N& e) {
        if (!e || !seen.insert(e.get()).second) return;
        switch (e->kind) {
            case K::Element:
            case K::SubVector:
            case K::ReadData:
            case K::Locate:
            case K::Length:
                if (computed(e->args[0])) {
                    for (size_t i = 1; i < e->args.size(); i++) visit(e->args[i]);
                    return;
                }
                break;
            case K::In:
                if (computed(e->args[1])) {
                    visit(e->args[0]);
                    return;
                }
                break;
            case K::Sym: {
                if (inputs.count(e->name)) return;
                auto t = tbl.find(e->name);
                if (t != tbl.end()) visit(t->second);
                return;
            }
            default:
                break;
        }
        for (auto& a : e->args) visit(a);
    };
    visit(func->args[0]);
    if (cands.empty()) return func;
    std::vector<N> in(func->args.begin() + 1, func->args.end());
    std::vector<std::pair<N, N>> main_tbl;
    for (auto& kv : func->tbl)
        if (!cand_set.count(kv.first->name)) main_tbl.push_back(kv);
    for (auto& c : cands) {
        N d = strip_initval(tbl[c]);
        Prelude p;
        N target = tsym[c];
        std::vector<std::pair<N, N>> sub_tbl = func->tbl;
        if (d->kind == K::BuildIdx) {
            p.index = true;
            N v = d->args[0];
            if (v->kind == K::Sym && tbl.count(v->name)) {
                target = v;
            } else {
                target = mk_sym(c + ".vec", v->type);
                sub_tbl.push_back({target, v});
            }
        }
        p.func = mk_func(func->name + "." + c, target, in, sub_tbl);
        pre.push_back(p);
        in.push_back(mk_sym(c, tsym[c]->type));
    }
    return mk_func(func->name, func->args[0], in, main_tbl);
}

std::vector<N> output_components(const N& func) {
    if (!func || func->kind != K::Func) throw IRError("expected a Func");
    std::map<std::string, N> tbl, tsym;
    for (auto& kv : func->tbl) {
        tbl[kv.first->name] = kv.second;
        tsym[kv.first->name] = kv.first;
    }
    std::vector<N> comps;
    std::set<std::string> seen;
    std::function<void(const N&)> walk = [&](const N& e) {
        if (e->kind == K::Sym) {
            auto t = tbl.find(e->name);
            if (t == tbl.end()) return;
            N d = strip_initval(t->second);
            if (d->kind == K::Reduce || d->kind == K::Op) {
                if (seen.insert(e->name).second) comps.push_back(tsym[e->name]);
                return;
            }
            walk(t->second);
            return;
        }
        for (auto& a : e->args) walk(a);
    };
    walk(func->args[0]);
    return comps;
}

std::vector<N> split_components(const N& func) {
    std::vector<N> in(func->args.begin() + 1, func->args.end());
    std::vector<N> out;
    for (const N& c : output_components(func))
        out.push_back(mk_func(func->name + "." + c->name, c, in, func->tbl));
    return out;
}

Plan lower_func(const N& func, bool vectorize) {
    if (!func || func->kind != K::Func) throw IRError("expected a Func");
    Lowerer L;
    L.func = func;
    L.vectorize = vectorize;
    return L.run();
}

static const char* aggop_name(AggOp o) {
    switch (o) {
        case AggOp::ADD: return "add";
        case AggOp::MUL: return "mul";
        case AggOp::MIN: return "min";
        case AggOp::MAX: return "max";
        case AggOp::AND: return "and";
        case AggOp::OR: return "or";
    }
    return "?";
}

std::string Plan::describe() const {
    std::ostringstream o;
    if (composite) {
        o << "func " << func_name << ": independent components, each a query of its own, "
          << "run one after another; output " << to_text(output) << "\n";
        return o.str();
    }
    if (source == MERGE)
        o << "func " << func_name << ": union merge of table inputs " << merge_a << " and "
          << merge_b << " (keys ascending"
          << (merge_off[0] || merge_off[1]
                  ? ", shifted by " + std::to_string(merge_off[0]) + " and " +
                        std::to_string(merge_off[1])
                  : std::string())
          << "), merge positions [" << to_text(start) << ", " << to_text(end) << ")\n";
    else
        o << "func " << func_name << ": scan "
          << (source == RANGE ? "range" : "table input " + std::to_string(driver))
          << " rows [" << to_text(start) << ", " << to_text(end) << ")\n";
    for (size_t j = 0; j < probes.size(); j++)
        o << "  probe " << j << ": table input " << probes[j].table << " key "
          << to_text(probes[j].key) << "\n";
    for (auto& r : reduces) {
        o << "  reduce " << r.sym_name << " -> " << r.state_type.str()
          << (r.sequential ? " (sequential fold, in row order)" : "") << "\n";
        if (r.sequential) {
            o << "    if " << (r.seq_pred ? to_text(r.seq_pred) : std::string("true")) << "\n";
            for (size_t f = 0; f < r.seq_next.size(); f++)
                o << "    field " << f << " " << r.fields[f].type.str() << " := "
                  << to_text(r.seq_next[f]) << "\n";
            continue;
        }
        for (size_t f = 0; f < r.fields.size(); f++) {
            auto& fl = r.fields[f];
            o << "    field " << f << " " << fl.type.str();
            if (fl.updated)
                o << " " << aggop_name(fl.op) << " if " << to_text(fl.pred)
                  << " of " << to_text(fl.contrib);
            else
                o << " (not updated)";
            o << "\n";
        }
    }
    auto seg_text = [&](const SegReducePlan& sp) {
        if (sp.ranged)
            o << "  ranged fold: for $key in [" << to_text(sp.lb) << ", " << to_text(sp.ub)
              << "), rows [" << to_text(sp.lo) << ", " << to_text(sp.hi) << ") of table input "
              << sp.table << "\n";
        else
            o << "  run reduction over the run-end column 0 of table input " << sp.table << "\n";
        for (size_t j = 0; j < sp.sub.probes.size(); j++)
            o << "    probe " << j << ": table input " << sp.sub.probes[j].table << " at key "
              << to_text(sp.sub.probes[j].key) << "\n";
        for (size_t f = 0; f < sp.fields.size(); f++) {
            auto& fl = sp.fields[f];
            o << "    run field " << f << " " << fl.type.str();
            if (fl.updated)
                o << " " << aggop_name(fl.op) << " if " << to_text(fl.pred) << " of "
                  << to_text(fl.contrib);
            else
                o << " (not updated)";
            o << "\n";
        }
        for (auto& v : sp.out_values) o << "    output " << to_text(v) << "\n";
    };
    if (seg) seg_text(*seg);
    if (runs) {
        o << "  run stage: the runs table " << runs->seg.vec_type.str() << " takes table input "
          << runs->table << "'s place\n";
        seg_text(runs->seg);
    }
    if (gj) {
        o << "  group join: build table input " << gj->build << ", probe table input "
          << gj->probe << "\n";
        o << "  build:\n" << gj->build_plan.describe();
        o << "  probe:\n" << gj->probe_plan.describe();
        for (size_t f = 0; f < gj->fields.size(); f++) {
            auto& g = gj->fields[f];
            o << "    group field " << f << " " << g.type.str();
            if (g.updated)
                o << " " << aggop_name(g.op) << " of "
                  << (g.const_contrib ? to_text(g.const_contrib) : "payload " + std::to_string(g.pay))
                  << (g.pred_pay >= 0 ? " if payload " + std::to_string(g.pred_pay) : "");
            else
                o << " (not updated)";
            o << "\n";
        }
        for (auto& v : gj->out_values) o << "    output " << to_text(v) << "\n";
    }
    for (auto& op : ops) {
        o << "  select " << op.sym_name << " -> " << op.vec_type.str() << " if "
          << to_text(op.pred) << "\n";
        for (auto& v : op.values) o << "    value " << to_text(v) << "\n";
    }
    return o.str();
}

}  // namespace rf
