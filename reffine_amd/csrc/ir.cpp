// Reffine IR nodes, typing and the S-expression reader.
#include "ir.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace rf {

namespace types {
const Type VOID_T(BT::VOID), BOOL_T(BT::BOOL), I8(BT::INT8), I16(BT::INT16),
    I32(BT::INT32), I64(BT::INT64), U8(BT::UINT8), U16(BT::UINT16),
    U32(BT::UINT32), U64(BT::UINT64), F32(BT::FLOAT32), F64(BT::FLOAT64),
    IDX_T(BT::IDX);
}

static void check(bool c, const std::string& msg) {
    if (!c) throw IRError(msg);
}

Type Type::valty() const {
    check(is_vector(), "valty of non-vector " + str());
    std::vector<Type> d(dtypes.begin() + 1, dtypes.end());
    std::vector<Enc> e;
    if (enc.size() > 1) e.assign(enc.begin() + 1, enc.end());
    if (dim == 1) return Type(BT::STRUCT, d);
    return Type(BT::VECTOR, d, dim - 1, e);
}

int Type::width() const {
    switch (bt) {
        case BT::BOOL: case BT::INT8: case BT::UINT8: return 1;
        case BT::INT16: case BT::UINT16: return 2;
        case BT::INT32: case BT::UINT32: case BT::FLOAT32: return 4;
        case BT::INT64: case BT::UINT64: case BT::FLOAT64: case BT::IDX:
            return 8;
        default: throw IRError("width of non-primitive " + str());
    }
}

// Same spelling as DataType::cppstr (type.h:212-249).
std::string Type::cstr() const {
    switch (bt) {
        case BT::VOID: return "void";
        case BT::BOOL: return "bool";
        case BT::INT8: return "int8_t";
        case BT::UINT8: return "uint8_t";
        case BT::INT16: return "int16_t";
        case BT::UINT16: return "uint16_t";
        case BT::INT32: return "int32_t";
        case BT::UINT32: return "uint32_t";
        case BT::INT64: case BT::IDX: return "int64_t";
        case BT::UINT64: return "uint64_t";
        case BT::FLOAT32: return "float";
        case BT::FLOAT64: return "double";
        default: throw IRError("no C spelling for " + str());
    }
}

std::string Type::str() const {
    switch (bt) {
        case BT::VOID: return "void";
        case BT::BOOL: return "b";
        case BT::INT8: return "i8";
        case BT::UINT8: return "u8";
        case BT::INT16: return "i16";
        case BT::UINT16: return "u16";
        case BT::INT32: return "i32";
        case BT::UINT32: return "u32";
        case BT::INT64: return "i64";
        case BT::UINT64: return "u64";
        case BT::FLOAT32: return "f32";
        case BT::FLOAT64: return "f64";
        case BT::IDX: return "x";
        case BT::STR: return "str";
        case BT::PTR: return "(ptr " + dtypes.at(0).str() + ")";
        case BT::STRUCT: {
            std::string s = "(struct";
            for (auto& d : dtypes) s += " " + d.str();
            return s + ")";
        }
        case BT::VECTOR: {
            std::string s = "(vec " + std::to_string(dim);
            for (size_t i = 0; i < dtypes.size(); i++) {
                bool ree = i < enc.size() && enc[i] == Enc::RUNEND;
                s += ree ? " (ree " + dtypes[i].str() + ")"
                         : " " + dtypes[i].str();
            }
            return s + ")";
        }
    }
    return "?";
}

static const char* kOpNames[] = {"add", "sub", "mul", "div", "max", "min",
                                  "mod", "sqrt", "pow", "abs", "neg", "ceil",
                                  "floor", "lt", "lte", "gt", "gte", "eq",
                                  "not", "and", "or"};

std::string mathop_name(MathOp op) { return kOpNames[(int)op]; }

// ------------------------------------------------------------ constructors

N mk_sym(const std::string& name, const Type& t) {
    auto n = std::make_shared<Node>(K::Sym, t);
    n->name = name;
    return n;
}

N mk_const(const Type& t, double v) {
    check(t.is_primitive(), "Const of non-primitive type " + t.str());
    auto n = std::make_shared<Node>(K::Const, t);
    n->val = v;
    return n;
}

N mk_cast(const Type& t, N a) {
    check(!t.is_struct() && !a->type.is_struct(), "Cast of struct");
    auto n = std::make_shared<Node>(K::Cast, t);
    n->args = {a};
    return n;
}

N mk_select(N c, N t, N f) {
    check(c->type == types::BOOL_T, "Select condition must be bool");
    check(t->type == f->type, "Select branches differ: " + t->type.str() +
                                  " vs " + f->type.str());
    auto n = std::make_shared<Node>(K::Select, t->type);
    n->args = {c, t, f};
    return n;
}

N mk_get(N v, int64_t col) {
    Type t;
    if (v->type.is_struct()) {
        check(col >= 0 && col < (int64_t)v->type.dtypes.size(),
              "Get column out of range");
        t = v->type.dtypes[col];
    } else {
        check(v->type.is_primitive() && col == 0, "Get on non-struct");
        t = v->type;
    }
    auto n = std::make_shared<Node>(K::Get, t);
    n->args = {v};
    n->idx = col;
    return n;
}

N mk_new(std::vector<N> vals) {
    std::vector<Type> d;
    for (auto& v : vals) d.push_back(v->type);
    check(!d.empty(), "New with no fields");
    auto n = std::make_shared<Node>(K::New, Type(BT::STRUCT, d));
    n->args = std::move(vals);
    return n;
}

N mk_nary(MathOp op, std::vector<N> args) {
    check(!args.empty(), "operator without operands");
    check(!args[0]->type.is_struct() && !args[0]->type.is_vector(),
          "arithmetic on struct/vector");
    Type t = args[0]->type;
    switch (op) {
        case MathOp::NOT:
        case MathOp::ABS: case MathOp::NEG: case MathOp::SQRT:
        case MathOp::CEIL: case MathOp::FLOOR:
            check(args.size() == 1, mathop_name(op) + " is unary");
            break;
        default:
            check(args.size() == 2, mathop_name(op) + " is binary");
            // expr.h:165-171: binary operands must have equal types.
            check(args[0]->type == args[1]->type,
                  mathop_name(op) + " operand types differ: " +
                      args[0]->type.str() + " vs " + args[1]->type.str());
    }
    switch (op) {
        case MathOp::LT: case MathOp::LTE: case MathOp::GT: case MathOp::GTE:
        case MathOp::EQ:
            t = types::BOOL_T;
            break;
        case MathOp::NOT: case MathOp::AND: case MathOp::OR:
            check(args[0]->type == types::BOOL_T,
                  mathop_name(op) + " needs bool operands");
            t = types::BOOL_T;
            break;
        case MathOp::MOD:
            check(!args[0]->type.is_float(), "mod of float");
            break;
        case MathOp::POW: case MathOp::CEIL: case MathOp::FLOOR:
            check(args[0]->type.is_float(), mathop_name(op) + " of non-float");
            break;
        default:
            break;
    }
    auto n = std::make_shared<Node>(K::Nary, t);
    n->op = op;
    n->args = std::move(args);
    return n;
}

N mk_element(N vec, N iter) {
    check(vec->type.is_vector(), "Element of non-vector");
    check(iter->type == vec->type.iterty(),
          "Element iterator type " + iter->type.str() + " vs " +
              vec->type.iterty().str());
    auto n = std::make_shared<Node>(K::Element, vec->type.valty());
    n->args = {vec, iter};
    return n;
}

N mk_in(N iter, N vec) {
    check(vec->type.is_vector(), "In over non-vector");
    check(iter->type == vec->type.iterty(), "In iterator type mismatch");
    auto n = std::make_shared<Node>(K::In, types::BOOL_T);
    n->args = {iter, vec};
    return n;
}

N mk_subvec(N vec, N start, N end) {
    check(vec->type.is_vector(), "SubVector of non-vector");
    check(start->type.is_idx() && end->type.is_idx(),
          "SubVector bounds must be idx");
    auto n = std::make_shared<Node>(K::SubVector, vec->type.valty());
    n->args = {vec, start, end};
    return n;
}

N mk_len(N vec, int64_t col) {
    check(vec->type.is_vector(), "Length of non-vector");
    check(col >= 0 && col < (int64_t)vec->type.dtypes.size(),
          "Length column out of range");
    auto n = std::make_shared<Node>(K::Length, types::IDX_T);
    n->args = {vec};
    n->idx = col;
    return n;
}

N mk_locate(N vec, N iter) {
    check(vec->type.is_vector(), "Locate in non-vector");
    check(vec->type.dim <= 2, "Locate in a vector of more than 2 dimensions");
    check(iter->type == vec->type.iterty(), "Locate iterator type mismatch");
    auto n = std::make_shared<Node>(K::Locate, types::IDX_T);
    n->args = {vec, iter};
    return n;
}

N mk_readdata(N vec, N idx, int64_t col) {
    check(vec->type.is_vector(), "ReadData of non-vector");
    check(col >= 0 && col < (int64_t)vec->type.dtypes.size(), "ReadData column out of range");
    check(idx->type.is_idx() || idx->type.is_int(), "ReadData index must be an integer");
    auto n = std::make_shared<Node>(K::ReadData, vec->type.dtypes[col]);
    n->args = {vec, idx};
    n->idx = col;
    return n;
}

N mk_initval(std::vector<N> inits, N val) {
    for (auto& i : inits) check(i->kind == K::Sym, "InitVal initialises symbols");
    auto n = std::make_shared<Node>(K::InitVal, val->type);
    n->iters = std::move(inits);
    n->args = {val};
    return n;
}

N mk_buildidx(N vec) {
    check(vec->type.is_vector(), "BuildIndex of non-vector");
    auto n = std::make_shared<Node>(K::BuildIdx, vec->type);
    n->args = {vec};
    return n;
}

N mk_op(std::vector<N> iters, N pred, std::vector<N> outputs) {
    std::vector<Type> d;
    std::vector<Enc> e;
    for (auto& it : iters) {
        check(it->kind == K::Sym && it->type.is_primitive(),
              "Op iterator must be a primitive symbol");
        d.push_back(it->type);
        e.push_back(Enc::FLAT);
    }
    for (auto& o : outputs) {
        check(o->type.is_val(), "Op output must be a value");
        d.push_back(o->type);
        e.push_back(Enc::FLAT);
    }
    check(pred->type == types::BOOL_T, "Op predicate must be bool");
    auto n = std::make_shared<Node>(
        K::Op, Type(BT::VECTOR, d, (int)iters.size(), e));
    n->iters = std::move(iters);
    n->args.push_back(pred);
    for (auto& o : outputs) n->args.push_back(o);
    return n;
}

N mk_reduce(N vec, N init, N state_sym, N val_sym, N acc) {
    check(vec->type.is_vector(), "Reduce over non-vector");
    check(state_sym->type == init->type, "Reduce state symbol type");
    check(val_sym->type == vec->type.rowty(), "Reduce value symbol type");
    check(acc->type == init->type, "Reduce accumulator changes state type");
    auto n = std::make_shared<Node>(K::Reduce, init->type);
    n->args = {vec, init, state_sym, val_sym, acc};
    return n;
}

N mk_func(const std::string& name, N output, std::vector<N> inputs,
          std::vector<std::pair<N, N>> tbl) {
    auto n = std::make_shared<Node>(K::Func, types::VOID_T);
    n->name = name;
    n->args.push_back(output);
    for (auto& i : inputs) n->args.push_back(i);
    n->tbl = std::move(tbl);
    return n;
}

// ------------------------------------------------------------- the reader

namespace {

struct SExpr {
    bool atom = false;
    std::string text;
    std::vector<SExpr> items;
};

struct Reader {
    const std::string& s;
    size_t p = 0;
    explicit Reader(const std::string& str) : s(str) {}
    void ws() {
        while (p < s.size()) {
            if (std::isspace((unsigned char)s[p])) {
                p++;
            } else if (s[p] == ';') {
                while (p < s.size() && s[p] != '\n') p++;
            } else {
                break;
            }
        }
    }
    SExpr read() {
        ws();
        check(p < s.size(), "IR text ended early");
        SExpr e;
        if (s[p] == '(') {
            p++;
            for (;;) {
                ws();
                check(p < s.size(), "unbalanced '(' in IR text");
                if (s[p] == ')') {
                    p++;
                    break;
                }
                e.items.push_back(read());
            }
        } else {
            check(s[p] != ')', "unexpected ')' in IR text");
            e.atom = true;
            size_t b = p;
            while (p < s.size() && !std::isspace((unsigned char)s[p]) &&
                   s[p] != '(' && s[p] != ')')
                p++;
            e.text = s.substr(b, p - b);
        }
        return e;
    }
};

struct Builder {
    std::map<std::string, N> syms;

    Type type(const SExpr& e) {
        if (e.atom) {
            static const std::map<std::string, BT> m = {
                {"void", BT::VOID}, {"b", BT::BOOL}, {"i8", BT::INT8},
                {"i16", BT::INT16}, {"i32", BT::INT32}, {"i64", BT::INT64},
                {"u8", BT::UINT8}, {"u16", BT::UINT16}, {"u32", BT::UINT32},
                {"u64", BT::UINT64}, {"f32", BT::FLOAT32},
                {"f64", BT::FLOAT64}, {"x", BT::IDX}};
            auto it = m.find(e.text);
            check(it != m.end(), "unknown type '" + e.text + "'");
            return Type(it->second);
        }
        check(!e.items.empty() && e.items[0].atom, "malformed type");
        const auto& h = e.items[0].text;
        if (h == "struct") {
            std::vector<Type> d;
            for (size_t i = 1; i < e.items.size(); i++) d.push_back(type(e.items[i]));
            check(!d.empty(), "empty struct type");
            return Type(BT::STRUCT, d);
        }
        if (h == "vec") {
            check(e.items.size() >= 3, "vector type needs a dim and columns");
            int dim = std::atoi(e.items[1].text.c_str());
            std::vector<Type> d;
            std::vector<Enc> en;
            for (size_t i = 2; i < e.items.size(); i++) {
                const auto& c = e.items[i];
                if (!c.atom && !c.items.empty() && c.items[0].atom &&
                    c.items[0].text == "ree") {
                    check(c.items.size() == 2, "malformed ree column");
                    d.push_back(type(c.items[1]));
                    en.push_back(Enc::RUNEND);
                } else {
                    d.push_back(type(c));
                    en.push_back(Enc::FLAT);
                }
            }
            check(dim > 0 && (int)d.size() >= dim, "vector dim > columns");
            return Type(BT::VECTOR, d, dim, en);
        }
        throw IRError("unknown type form '" + h + "'");
    }

    N sym(const SExpr& e) {
        check(!e.atom && e.items.size() == 3 && e.items[0].text == "sym",
              "expected (sym NAME TYPE)");
        const auto& name = e.items[1].text;
        Type t = type(e.items[2]);
        auto it = syms.find(name);
        if (it != syms.end()) {
            check(it->second->type == t, "symbol " + name + " retyped");
            return it->second;
        }
        auto s = mk_sym(name, t);
        syms[name] = s;
        return s;
    }

    int64_t integer(const SExpr& e) {
        check(e.atom, "expected an integer");
        char* end = nullptr;
        long long v = std::strtoll(e.text.c_str(), &end, 10);
        check(end && *end == 0, "bad integer '" + e.text + "'");
        return v;
    }

    N expr(const SExpr& e) {
        check(!e.atom && !e.items.empty() && e.items[0].atom,
              "expected an expression list");
        const auto& h = e.items[0].text;
        auto arg = [&](size_t i) {
            check(i < e.items.size(), "missing operand of " + h);
            return expr(e.items[i]);
        };
        if (h == "sym") return sym(e);
        if (h == "const") {
            check(e.items.size() == 3, "(const TYPE VALUE)");
            Type t = type(e.items[1]);
            const auto& v = e.items[2].text;
            double d;
            if (v == "inf") d = INFINITY;
            else if (v == "-inf") d = -INFINITY;
            else if (v == "nan") d = NAN;
            else {
                char* end = nullptr;
                d = std::strtod(v.c_str(), &end);
                check(end && *end == 0, "bad constant '" + v + "'");
            }
            return mk_const(t, d);
        }
        if (h == "cast") return mk_cast(type(e.items.at(1)), arg(2));
        if (h == "sel") return mk_select(arg(1), arg(2), arg(3));
        if (h == "get") return mk_get(arg(1), integer(e.items.at(2)));
        if (h == "new") {
            std::vector<N> v;
            for (size_t i = 1; i < e.items.size(); i++) v.push_back(arg(i));
            return mk_new(v);
        }
        if (h == "elem") return mk_element(arg(1), arg(2));
        if (h == "in") return mk_in(arg(1), arg(2));
        if (h == "subvec") return mk_subvec(arg(1), arg(2), arg(3));
        if (h == "len") return mk_len(arg(1), integer(e.items.at(2)));
        if (h == "locate") return mk_locate(arg(1), arg(2));
        if (h == "buildidx") return mk_buildidx(arg(1));
        if (h == "initval") {
            check(e.items.size() == 3, "(initval (SYMS) VAL)");
            std::vector<N> inits;
            for (auto& i : e.items[1].items) inits.push_back(sym(i));
            return mk_initval(inits, expr(e.items[2]));
        }
        if (h == "readdata") return mk_readdata(arg(1), arg(2), integer(e.items.at(3)));
        if (h == "op") {
            check(e.items.size() == 4, "(op (ITERS) PRED (OUTPUTS))");
            std::vector<N> iters, outs;
            for (auto& i : e.items[1].items) iters.push_back(sym(i));
            N pred = expr(e.items[2]);
            for (auto& o : e.items[3].items) outs.push_back(expr(o));
            return mk_op(iters, pred, outs);
        }
        if (h == "reduce") {
            check(e.items.size() == 6, "(reduce VEC INIT S V ACC)");
            return mk_reduce(arg(1), arg(2), sym(e.items[3]), sym(e.items[4]),
                             arg(5));
        }
        for (int i = 0; i <= (int)MathOp::OR; i++) {
            if (h == kOpNames[i]) {
                std::vector<N> v;
                for (size_t j = 1; j < e.items.size(); j++) v.push_back(arg(j));
                return mk_nary((MathOp)i, v);
            }
        }
        throw IRError("unknown IR node '" + h + "'");
    }

    N func(const SExpr& e) {
        check(!e.atom && e.items.size() == 5 && e.items[0].text == "func",
              "expected (func NAME OUTPUT (INPUTS) (TABLE))");
        std::string name = e.items[1].text;
        N out = sym(e.items[2]);
        std::vector<N> inputs;
        for (auto& i : e.items[3].items) inputs.push_back(sym(i));
        std::vector<std::pair<N, N>> tbl;
        for (auto& d : e.items[4].items) {
            check(!d.atom && d.items.size() == 2, "table entry (SYM EXPR)");
            N s = sym(d.items[0]);
            N v = expr(d.items[1]);
            check(s->type == v->type, "table entry type mismatch for " +
                                          s->name);
            tbl.push_back({s, v});
        }
        return mk_func(name, out, inputs, tbl);
    }
};

}  // namespace

N parse_func(const std::string& text) {
    Reader r(text);
    SExpr e = r.read();
    r.ws();
    check(r.p == text.size(), "trailing text after IR function");
    Builder b;
    return b.func(e);
}

static std::string num(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

std::string to_text(const N& n) {
    switch (n->kind) {
        case K::Sym: return "(sym " + n->name + " " + n->type.str() + ")";
        case K::Const: return "(const " + n->type.str() + " " + num(n->val) + ")";
        case K::Cast:
            return "(cast " + n->type.str() + " " + to_text(n->args[0]) + ")";
        case K::Select:
            return "(sel " + to_text(n->args[0]) + " " + to_text(n->args[1]) +
                   " " + to_text(n->args[2]) + ")";
        case K::Get:
            return "(get " + to_text(n->args[0]) + " " + std::to_string(n->idx) + ")";
        case K::New: {
            std::string s = "(new";
            for (auto& a : n->args) s += " " + to_text(a);
            return s + ")";
        }
        case K::Nary: {
            std::string s = "(" + mathop_name(n->op);
            for (auto& a : n->args) s += " " + to_text(a);
            return s + ")";
        }
        case K::Element:
            return "(elem " + to_text(n->args[0]) + " " + to_text(n->args[1]) + ")";
        case K::In:
            return "(in " + to_text(n->args[0]) + " " + to_text(n->args[1]) + ")";
        case K::SubVector:
            return "(subvec " + to_text(n->args[0]) + " " + to_text(n->args[1]) +
                   " " + to_text(n->args[2]) + ")";
        case K::Length:
            return "(len " + to_text(n->args[0]) + " " + std::to_string(n->idx) + ")";
        case K::Locate:
            return "(locate " + to_text(n->args[0]) + " " + to_text(n->args[1]) + ")";
        case K::BuildIdx:
            return "(buildidx " + to_text(n->args[0]) + ")";
        case K::InitVal: {
            std::string s = "(initval (";
            for (size_t i = 0; i < n->iters.size(); i++)
                s += (i ? " " : "") + to_text(n->iters[i]);
            return s + ") " + to_text(n->args[0]) + ")";
        }
        case K::ReadData:
            return "(readdata " + to_text(n->args[0]) + " " + to_text(n->args[1]) + " " +
                   std::to_string(n->idx) + ")";
        case K::Op: {
            std::string s = "(op (";
            for (size_t i = 0; i < n->iters.size(); i++)
                s += (i ? " " : "") + to_text(n->iters[i]);
            s += ") " + to_text(n->args[0]) + " (";
            for (size_t i = 1; i < n->args.size(); i++)
                s += (i > 1 ? " " : "") + to_text(n->args[i]);
            return s + "))";
        }
        case K::Reduce: {
            std::string s = "(reduce";
            for (auto& a : n->args) s += " " + to_text(a);
            return s + ")";
        }
        case K::Func: {
            std::string s = "(func " + n->name + " " + to_text(n->args[0]) + " (";
            for (size_t i = 1; i < n->args.size(); i++)
                s += (i > 1 ? " " : "") + to_text(n->args[i]);
            s += ") (";
            for (size_t i = 0; i < n->tbl.size(); i++)
                s += (i ? " (" : "(") + to_text(n->tbl[i].first) + " " +
                     to_text(n->tbl[i].second) + ")";
            return s + "))";
        }
    }
    return "?";
}

}  // namespace rf
