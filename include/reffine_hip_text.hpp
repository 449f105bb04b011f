// reffine_hip_text.hpp -- the reffine-side half of the drop-in boundary.
//
// Header-only C++17 over reffine's own IR classes (include/reffine/ir/*.h,
// include/reffine/pass/base/visitor.h of ampersand-projects/reffine): it
// writes a reffine Func as the IR text that rf_compile (reffine_hip.h)
// reads, one S-expression per IR node, and wraps rf_compile so that a
// reffine program replaces
//
//     auto fn = compile_op<void (*)(long*, ArrowTable*)>(func, vectorize);
//     fn(&out, tbl);                                     // utils.h:47-54
//
// with
//
//     reffine_hip::HipQuery q(func);                     // to_rf_text + rf_compile
//     q(&out, {reffine_hip::table(tbl)});                // rf_query_run
//
// Text grammar (read by the library's parser):
//   type  := void | b | i8 | i16 | i32 | i64 | u8 | u16 | u32 | u64 | f32 | f64 | x
//          | (struct TYPE...) | (vec DIM COL...)   COL := TYPE | (ree TYPE)
//   expr  := (sym NAME#ID TYPE) | (const TYPE VALUE) | (cast TYPE E) | (sel C T F)
//          | (get E COL) | (new E...) | (OPNAME E...) | (elem VEC IT) | (in IT VEC)
//          | (subvec VEC START END) | (len VEC COL) | (locate VEC IT)
//          | (readdata VEC IDX COL) | (initval (SYM...) E) | (buildidx VEC)
//          | (op (SYM...) PRED (E...)) | (reduce VEC INIT STATE VAL ACC)
//   func  := (func NAME OUTSYM (INPUTSYM...) ((SYM E)...))
// A Reduce's init and acc lambdas (include/reffine/ir/op.h:97-114) are applied
// to two fresh symbols, as the Reduce constructor does for its type check.
// Symbols are told apart by identity (NAME#ID, ID per SymNode object).
#ifndef REFFINE_HIP_TEXT_HPP_
#define REFFINE_HIP_TEXT_HPP_

#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "reffine/arrow/abi.h"
#include "reffine/ir/expr.h"
#include "reffine/ir/op.h"
#include "reffine/ir/op_to_loop.h"
#include "reffine/ir/stmt.h"
#include "reffine/pass/base/visitor.h"
#include "reffine_hip.h"

namespace reffine_hip {

class TextWriter : public reffine::Visitor {
public:
    std::string write(const std::shared_ptr<reffine::Func>& fn)
    {
        os_.str("");
        ids_.clear();
        fn->Accept(*this);
        return os_.str();
    }

    static std::string type_text(const reffine::DataType& t)
    {
        using reffine::BaseType;
        switch (t.btype) {
            case BaseType::VOID: return "void";
            case BaseType::BOOL: return "b";
            case BaseType::INT8: return "i8";
            case BaseType::INT16: return "i16";
            case BaseType::INT32: return "i32";
            case BaseType::INT64: return "i64";
            case BaseType::UINT8: return "u8";
            case BaseType::UINT16: return "u16";
            case BaseType::UINT32: return "u32";
            case BaseType::UINT64: return "u64";
            case BaseType::FLOAT32: return "f32";
            case BaseType::FLOAT64: return "f64";
            case BaseType::IDX: return "x";
            case BaseType::STRUCT: {
                std::string s = "(struct";
                for (const auto& d : t.dtypes) s += " " + type_text(d);
                return s + ")";
            }
            case BaseType::VECTOR: {
                std::string s = "(vec " + std::to_string(t.dim);
                for (size_t i = 0; i < t.dtypes.size(); i++) {
                    const bool ree = i < t.encodings.size() &&
                                     t.encodings[i] == reffine::EncodeType::RUNEND;
                    s += ree ? " (ree " + type_text(t.dtypes[i]) + ")"
                             : " " + type_text(t.dtypes[i]);
                }
                return s + ")";
            }
            default:
                throw std::runtime_error("reffine_hip: no IR text for type " + t.str());
        }
    }

private:
    std::ostringstream os_;
    std::map<const reffine::SymNode*, int> ids_;

    void emit(const reffine::Expr& e) { e->Accept(*this); }

    void emit_all(const std::vector<reffine::Expr>& v)
    {
        for (size_t i = 0; i < v.size(); i++) {
            if (i) os_ << " ";
            emit(v[i]);
        }
    }

    void emit_syms(const std::vector<reffine::Sym>& v)
    {
        os_ << "(";
        for (size_t i = 0; i < v.size(); i++) {
            if (i) os_ << " ";
            emit(v[i]);
        }
        os_ << ")";
    }

    // NAME#ID: the name with the grammar's delimiters replaced, and an id
    // per SymNode object.
    std::string ref(const reffine::SymNode& s)
    {
        auto it = ids_.find(&s);
        int id;
        if (it != ids_.end()) {
            id = it->second;
        } else {
            id = (int)ids_.size() + 1;
            ids_.emplace(&s, id);
        }
        std::string name = s.name.empty() ? "s" : s.name;
        for (char& c : name)
            if (c == ' ' || c == '(' || c == ')' || c == '#' || c == ';') c = '_';
        return name + "#" + std::to_string(id);
    }

    static const char* op_name(reffine::MathOp op)
    {
        using reffine::MathOp;
        switch (op) {
            case MathOp::ADD: return "add";
            case MathOp::SUB: return "sub";
            case MathOp::MUL: return "mul";
            case MathOp::DIV: return "div";
            case MathOp::MAX: return "max";
            case MathOp::MIN: return "min";
            case MathOp::MOD: return "mod";
            case MathOp::SQRT: return "sqrt";
            case MathOp::POW: return "pow";
            case MathOp::ABS: return "abs";
            case MathOp::NEG: return "neg";
            case MathOp::CEIL: return "ceil";
            case MathOp::FLOOR: return "floor";
            case MathOp::LT: return "lt";
            case MathOp::LTE: return "lte";
            case MathOp::GT: return "gt";
            case MathOp::GTE: return "gte";
            case MathOp::EQ: return "eq";
            case MathOp::NOT: return "not";
            case MathOp::AND: return "and";
            case MathOp::OR: return "or";
            default:
                throw std::runtime_error(
                    "reffine_hip: predicate operators (implies, iff, forall, exists) "
                    "belong to the solver, not to a query");
        }
    }

    void Visit(reffine::SymNode& s) final
    {
        os_ << "(sym " << ref(s) << " " << type_text(s.type) << ")";
    }

    void Visit(reffine::Func& f) final
    {
        // The output must be a symbol of the table; any other expression is
        // named here.
        auto out = std::dynamic_pointer_cast<reffine::SymNode>(f.output);
        reffine::SymTable tbl = f.tbl;
        if (!out) {
            out = std::make_shared<reffine::SymNode>("out", f.output->type);
            tbl[out] = f.output;
        }
        std::string name = f.name;
        for (char& c : name)
            if (c == ' ' || c == '(' || c == ')') c = '_';
        os_ << "(func " << name << " ";
        emit(out);
        os_ << " ";
        emit_syms(f.inputs);
        os_ << " (";
        bool first = true;
        for (auto& kv : tbl) {
            os_ << (first ? "(" : " (");
            first = false;
            emit(kv.first);
            os_ << " ";
            emit(kv.second);
            os_ << ")";
        }
        os_ << "))";
    }

    void Visit(reffine::Const& c) final
    {
        os_ << "(const " << type_text(c.type) << " ";
        if (std::isnan(c.val)) {
            os_ << "nan";
        } else if (std::isinf(c.val)) {
            os_ << (c.val > 0 ? "inf" : "-inf");
        } else {
            char buf[40];
            std::snprintf(buf, sizeof buf, "%.17g", c.val);  // round-trips exactly
            os_ << buf;
        }
        os_ << ")";
    }

    void Visit(reffine::Cast& c) final
    {
        os_ << "(cast " << type_text(c.type) << " ";
        emit(c.arg);
        os_ << ")";
    }

    void Visit(reffine::Select& s) final
    {
        os_ << "(sel ";
        emit(s.cond);
        os_ << " ";
        emit(s.true_body);
        os_ << " ";
        emit(s.false_body);
        os_ << ")";
    }

    void Visit(reffine::Get& g) final
    {
        os_ << "(get ";
        emit(g.val);
        os_ << " " << g.col << ")";
    }

    void Visit(reffine::New& n) final
    {
        os_ << "(new ";
        emit_all(n.vals);
        os_ << ")";
    }

    void Visit(reffine::NaryExpr& e) final
    {
        os_ << "(" << op_name(e.op) << " ";
        emit_all(e.args);
        os_ << ")";
    }

    void Visit(reffine::Element& e) final
    {
        os_ << "(elem ";
        emit(e.vec);
        os_ << " ";
        emit(e.iter);
        os_ << ")";
    }

    void Visit(reffine::In& e) final
    {
        os_ << "(in ";
        emit(e.iter);
        os_ << " ";
        emit(e.vec);
        os_ << ")";
    }

    void Visit(reffine::Op& op) final
    {
        os_ << "(op ";
        emit_syms(op.iters);
        os_ << " ";
        emit(op.pred);
        os_ << " (";
        emit_all(op.outputs);
        os_ << "))";
    }

    void Visit(reffine::Reduce& r) final
    {
        const reffine::Expr init = r.init();
        auto state = std::make_shared<reffine::SymNode>("state", init->type);
        auto val = std::make_shared<reffine::SymNode>("val", r.vec->type.rowty());
        const reffine::Expr acc = r.acc(state, val);
        os_ << "(reduce ";
        emit(r.vec);
        os_ << " ";
        emit(init);
        os_ << " ";
        emit(state);
        os_ << " ";
        emit(val);
        os_ << " ";
        emit(acc);
        os_ << ")";
    }

    void Visit(reffine::SubVector& s) final
    {
        os_ << "(subvec ";
        emit(s.vec);
        os_ << " ";
        emit(s.start);
        os_ << " ";
        emit(s.end);
        os_ << ")";
    }

    void Visit(reffine::Length& l) final
    {
        os_ << "(len ";
        emit(l.vec);
        os_ << " " << l.col << ")";
    }

    void Visit(reffine::ReadData& r) final
    {
        os_ << "(readdata ";
        emit(r.vec);
        os_ << " ";
        emit(r.idx);
        os_ << " " << r.col << ")";
    }

    void Visit(reffine::InitVal& v) final
    {
        os_ << "(initval ";
        emit_syms(v.inits);
        os_ << " ";
        emit(v.val);
        os_ << ")";
    }

    // Locate and BuildIndex are Calls (op_to_loop.h:106-117, 184-190).
    void Visit(reffine::Call& c) final
    {
        if (c.name == "vector_locate" && c.args.size() == 2) {
            os_ << "(locate ";
            emit(c.args[0]);
            os_ << " ";
            emit(c.args[1]);
            os_ << ")";
        } else if (c.name == "build_vector_index" && c.args.size() == 1) {
            os_ << "(buildidx ";
            emit(c.args[0]);
            os_ << ")";
        } else {
            throw std::runtime_error("reffine_hip: call '" + c.name +
                                     "' is loop-level IR, not a query node");
        }
    }
};

// The IR text of `fn` for rf_compile.
inline std::string to_rf_text(const std::shared_ptr<reffine::Func>& fn)
{
    return TextWriter().write(fn);
}

// An rf_arg for a table argument: a reffine ArrowTable (include/reffine/arrow/
// abi.h:56-69) uploaded to the device once per call of `table`; the handle
// is owned by the returned Table.
class Table {
public:
    explicit Table(::ArrowTable* t)
    {
        if (rf_table_upload(t->schema, t->array, t->dim, &h_) != RF_OK)
            throw std::runtime_error(rf_last_error());
    }
    ~Table() { if (h_) rf_table_free(h_); }
    Table(const Table&) = delete;
    Table& operator=(const Table&) = delete;
    rf_table* handle() const { return h_; }

private:
    rf_table* h_ = nullptr;
};

inline rf_arg table(const Table& t)
{
    rf_arg a{};
    a.kind = 0;
    a.u.table = t.handle();
    return a;
}

inline rf_arg i64(int64_t v)
{
    rf_arg a{};
    a.kind = 1;
    a.u.i64 = v;
    return a;
}

inline rf_arg f64(double v)
{
    rf_arg a{};
    a.kind = 2;
    a.u.f64 = v;
    return a;
}

// compile_op's counterpart: the Func compiled for the GPU; calling it runs
// the generated entry point's work on device tables.
class HipQuery {
public:
    explicit HipQuery(const std::shared_ptr<reffine::Func>& fn, bool deterministic = false)
    {
        const std::string text = to_rf_text(fn);
        if (rf_compile(text.c_str(), deterministic ? RF_COMPILE_DETERMINISTIC : 0u, &q_) != RF_OK)
            throw std::runtime_error(rf_last_error());
    }
    ~HipQuery() { if (q_) rf_query_free(q_); }
    HipQuery(const HipQuery&) = delete;
    HipQuery& operator=(const HipQuery&) = delete;

    // `void name(OutT* output_addr, In0 a0, ...)`: a reduction's struct goes
    // to output_addr; an Op's table handle to *out_table (rf_table_free it).
    void operator()(void* output_addr, const std::vector<rf_arg>& args,
                    rf_table** out_table = nullptr) const
    {
        if (rf_query_run(q_, args.data(), (int64_t)args.size(), output_addr, out_table) != RF_OK)
            throw std::runtime_error(rf_last_error());
    }

    rf_query* handle() const { return q_; }

private:
    rf_query* q_ = nullptr;
};

}  // namespace reffine_hip

#endif  // REFFINE_HIP_TEXT_HPP_
