// Reffine IR as seen by the GPU lowering.
//
// The node set follows the reffine IR (include/reffine/ir/expr.h,
// include/reffine/ir/op.h, include/reffine/ir/op_to_loop.h,
// include/reffine/ir/stmt.h) and its type system
// (include/reffine/base/type.h:17-278). Nodes are a single tagged struct so
// the lowering can pattern-match and rewrite them without a visitor
// hierarchy. Typing rules are enforced at construction exactly where the
// reference asserts (binary operands of equal type, expr.h:165-171; Select
// branch types, expr.h:44-52; Element/In iterator types, op.h:55-92).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rf {

struct IRError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// include/reffine/base/type.h:17-42
enum class BT {
    VOID, BOOL, INT8, INT16, INT32, INT64, UINT8, UINT16, UINT32, UINT64,
    FLOAT32, FLOAT64, STRUCT, PTR, STR, IDX, VECTOR
};
enum class Enc { FLAT, RUNEND };

struct Type {
    BT bt = BT::VOID;
    std::vector<Type> dtypes;
    int dim = 0;
    std::vector<Enc> enc;

    Type() = default;
    explicit Type(BT b) : bt(b) {}
    Type(BT b, std::vector<Type> d, int dm = 0, std::vector<Enc> e = {})
        : bt(b), dtypes(std::move(d)), dim(dm), enc(std::move(e)) {}

    bool operator==(const Type& o) const {
        return bt == o.bt && dtypes == o.dtypes && dim == o.dim;
    }
    bool operator!=(const Type& o) const { return !(*this == o); }
    bool is_vector() const { return bt == BT::VECTOR; }
    bool is_struct() const { return bt == BT::STRUCT; }
    bool is_idx() const { return bt == BT::IDX; }
    bool is_float() const { return bt == BT::FLOAT32 || bt == BT::FLOAT64; }
    bool is_int() const {
        return bt == BT::INT8 || bt == BT::INT16 || bt == BT::INT32 ||
               bt == BT::INT64 || bt == BT::UINT8 || bt == BT::UINT16 ||
               bt == BT::UINT32 || bt == BT::UINT64;
    }
    bool is_unsigned() const {
        return bt == BT::UINT8 || bt == BT::UINT16 || bt == BT::UINT32 ||
               bt == BT::UINT64;
    }
    bool is_primitive() const {
        return is_int() || is_float() || is_idx() || bt == BT::BOOL;
    }
    bool is_val() const { return !is_vector(); }
    // type.h:132-142: drops the leading iterator column.
    Type valty() const;
    Type iterty() const { return dtypes.at(0); }
    // type.h:150-154
    Type rowty() const { return Type(BT::STRUCT, dtypes); }
    int width() const;           // bytes of a primitive
    std::string cstr() const;    // C spelling used by the emitters
    std::string str() const;     // IR text spelling
};

namespace types {
extern const Type VOID_T, BOOL_T, I8, I16, I32, I64, U8, U16, U32, U64, F32,
    F64, IDX_T;
}

// include/reffine/base/type.h:252-278
enum class MathOp {
    ADD, SUB, MUL, DIV, MAX, MIN, MOD, SQRT, POW, ABS, NEG, CEIL, FLOOR,
    LT, LTE, GT, GTE, EQ, NOT, AND, OR
};

enum class K {
    Sym, Const, Cast, Select, Get, New, Nary, Element, In, SubVector,
    Length, Op, Reduce, Func, Locate, ReadData, InitVal, BuildIdx
};

struct Node;
using N = std::shared_ptr<Node>;

struct Node {
    K kind;
    Type type;
    MathOp op = MathOp::ADD;
    std::vector<N> args;   // operands (meaning depends on kind)
    std::string name;      // Sym / Func name
    double val = 0;        // Const (reffine stores every constant as double)
    int64_t idx = 0;       // Get column, Length column
    // Op: iters + pred(args[0]) + outputs(args[1..]); Reduce: args =
    // {vec, init, state_sym, val_sym, acc}; Func: args = {output, inputs...}
    std::vector<N> iters;
    std::vector<std::pair<N, N>> tbl;  // Func symbol table (sym -> expr)

    Node(K k, Type t) : kind(k), type(std::move(t)) {}
};

// Constructors with the reference's typing checks.
N mk_sym(const std::string& name, const Type& t);
N mk_const(const Type& t, double v);
N mk_cast(const Type& t, N a);
N mk_select(N c, N t, N f);
N mk_get(N v, int64_t col);
N mk_new(std::vector<N> vals);
N mk_nary(MathOp op, std::vector<N> args);
N mk_element(N vec, N iter);
N mk_in(N iter, N vec);
N mk_subvec(N vec, N start, N end);
N mk_len(N vec, int64_t col);
// include/reffine/ir/op_to_loop.h:106-117: the row of `iter` in vec's index
// (idx, -1 when absent).
N mk_locate(N vec, N iter);
// include/reffine/ir/op_to_loop.h:27-41: column `col` of vec at row `idx`.
N mk_readdata(N vec, N idx, int64_t col);
// include/reffine/ir/expr.h:16-26: `val`, with the symbols `inits` computed
// first (iters holds them).
N mk_initval(std::vector<N> inits, N val);
// include/reffine/ir/op_to_loop.h:184-190: vec with an index built over it.
N mk_buildidx(N vec);
N mk_op(std::vector<N> iters, N pred, std::vector<N> outputs);
N mk_reduce(N vec, N init, N state_sym, N val_sym, N acc);
N mk_func(const std::string& name, N output, std::vector<N> inputs,
          std::vector<std::pair<N, N>> tbl);

// Parses the S-expression IR text produced by reffine_amd.reffiner.
N parse_func(const std::string& text);

std::string mathop_name(MathOp op);
std::string to_text(const N& n);  // debug / cache key printer

}  // namespace rf
