"""Reffine IR builder DSL, mirrored in Python.

Names, argument meaning and typing rules follow the reference's builder
(include/reffine/builder/reffiner.h:13-213), IR nodes
(include/reffine/ir/expr.h, op.h, op_to_loop.h, stmt.h) and type system
(include/reffine/base/type.h:17-278):

  * `_expr` operators build Add/Sub/Mul/Div/Mod/Lt/Lte/Gt/Gte/Equals/And/Or/
    Not nodes; binary operands must have equal types (expr.h:165-171);
  * `v[k]` with an int is Get (column k of a struct), `vec[t]` with an
    expression is Element;
  * `_red(vec, init_fn, acc_fn)` evaluates the lambdas on symbols, as the
    reference's Reduce constructor does (op.h:97-114);
  * `_func(name, out_sym, inputs)` with `fn.tbl[sym] = expr` (stmt.h:8-27).

`Func.ir()` writes the S-expression text that the C ABI's rf_compile reads.
"""
from __future__ import annotations

import itertools
from typing import Callable, Sequence

# --------------------------------------------------------------------- types


class DataType:
    """include/reffine/base/type.h:45-250."""

    _PRIM = ("void", "b", "i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64",
             "f32", "f64", "x")

    def __init__(self, btype: str, dtypes: Sequence["DataType"] = (), dim: int = 0,
                 encodings: Sequence[str] = ()):
        self.btype = btype
        self.dtypes = tuple(dtypes)
        self.dim = dim
        self.encodings = tuple(encodings) if encodings else tuple("flat" for _ in dtypes)
        if btype == "struct":
            assert dim == 0 and len(self.dtypes) > 0
        elif btype == "vec":
            assert dim > 0 and len(self.dtypes) >= dim
            assert len(self.dtypes) == len(self.encodings)
        else:
            assert btype in self._PRIM, btype
            assert not self.dtypes and dim == 0

    def __eq__(self, o):  # type.h:64-68 (encodings are not compared)
        return (isinstance(o, DataType) and self.btype == o.btype and
                self.dtypes == o.dtypes and self.dim == o.dim)

    def __hash__(self):
        return hash((self.btype, self.dtypes, self.dim))

    def is_vector(self):
        return self.btype == "vec"

    def is_struct(self):
        return self.btype == "struct"

    def is_idx(self):
        return self.btype == "x"

    def is_float(self):
        return self.btype in ("f32", "f64")

    def is_int(self):
        return self.btype in ("i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64")

    def is_primitive(self):
        return self.is_int() or self.is_float() or self.is_idx() or self.btype == "b"

    def is_val(self):
        return not self.is_vector()

    def valty(self) -> "DataType":  # type.h:132-142
        assert self.is_vector()
        if self.dim == 1:
            return DataType("struct", self.dtypes[1:])
        return DataType("vec", self.dtypes[1:], self.dim - 1, self.encodings[1:])

    def iterty(self) -> "DataType":
        assert self.is_vector()
        return self.dtypes[0]

    def rowty(self) -> "DataType":  # type.h:150-154
        assert self.is_vector()
        return DataType("struct", self.dtypes)

    def ir(self) -> str:
        if self.btype == "struct":
            return "(struct " + " ".join(d.ir() for d in self.dtypes) + ")"
        if self.btype == "vec":
            cols = []
            for d, e in zip(self.dtypes, self.encodings):
                cols.append(f"(ree {d.ir()})" if e == "runend" else d.ir())
            return f"(vec {self.dim} " + " ".join(cols) + ")"
        return self.btype

    def __repr__(self):
        return self.ir()


_i8_t = DataType("i8")
_i16_t = DataType("i16")
_i32_t = DataType("i32")
_i64_t = DataType("i64")
_u8_t = DataType("u8")
_u16_t = DataType("u16")
_u32_t = DataType("u32")
_u64_t = DataType("u64")
_f32_t = DataType("f32")
_f64_t = DataType("f64")
_ch_t = _i8_t
_idx_t = DataType("x")
_bool_t = DataType("b")
_void_t = DataType("void")


def _vec_t(dim: int, *dtypes: DataType, encodings: Sequence[str] = ()) -> DataType:
    return DataType("vec", dtypes, dim, encodings or ["flat"] * len(dtypes))


def _struct_t(*dtypes: DataType) -> DataType:
    return DataType("struct", dtypes)


# ---------------------------------------------------------------------- nodes

_ids = itertools.count()


def _check(c: bool, msg: str):
    if not c:
        raise TypeError(msg)


class Expr:
    """Base of every IR node (include/reffine/ir/node.h:20-51)."""

    def __init__(self, type_: DataType):
        self.type = type_

    # operators (reffiner.h:35-70)
    def __add__(self, o): return _add(self, o)
    def __sub__(self, o): return _sub(self, o)
    def __mul__(self, o): return _mul(self, o)
    def __truediv__(self, o): return _div(self, o)
    def __mod__(self, o): return _mod(self, o)
    def __neg__(self): return _neg(self)
    def __lt__(self, o): return _lt(self, o)
    def __le__(self, o): return _lte(self, o)
    def __gt__(self, o): return _gt(self, o)
    def __ge__(self, o): return _gte(self, o)
    def __eq__(self, o): return _eq(self, o)  # type: ignore[override]
    def __ne__(self, o): return _not(_eq(self, o))  # type: ignore[override]
    def __invert__(self): return _not(self)
    def __and__(self, o): return _and(self, o)
    def __or__(self, o): return _or(self, o)
    __hash__ = object.__hash__

    def __getitem__(self, k):
        if isinstance(k, int):
            return _get(self, k)
        if isinstance(k, (list, tuple)) and len(k) == 1:
            k = k[0]
        return _elem(self, k)

    def ir(self) -> str:
        raise NotImplementedError


class Sym(Expr):
    def __init__(self, name: str, t):
        super().__init__(t if isinstance(t, DataType) else t.type)
        _check(all(c not in name for c in " ()#;"), f"bad symbol name {name!r}")
        self.name = name
        self.uid = next(_ids)

    def ref(self) -> str:
        return f"{self.name}#{self.uid}"

    def ir(self) -> str:
        return f"(sym {self.ref()} {self.type.ir()})"


class Const(Expr):
    """Every constant value is held as a double (expr.h:76-85)."""

    def __init__(self, t: DataType, val):
        _check(t.is_primitive(), "Const of non-primitive type")
        super().__init__(t)
        self.val = float(val)

    def ir(self):
        v = self.val
        if v != v:
            s = "nan"
        elif v in (float("inf"), float("-inf")):
            s = "inf" if v > 0 else "-inf"
        else:
            s = repr(v)
        return f"(const {self.type.ir()} {s})"


class Cast(Expr):
    def __init__(self, t: DataType, arg: Expr):
        _check(not t.is_struct() and not arg.type.is_struct(), "Cast of struct")
        super().__init__(t)
        self.arg = arg

    def ir(self):
        return f"(cast {self.type.ir()} {self.arg.ir()})"


class Select(Expr):
    def __init__(self, cond: Expr, t: Expr, f: Expr):
        _check(cond.type == _bool_t, "Select condition must be bool")
        _check(t.type == f.type, f"Select branches differ: {t.type} vs {f.type}")
        super().__init__(t.type)
        self.cond, self.t, self.f = cond, t, f

    def ir(self):
        return f"(sel {self.cond.ir()} {self.t.ir()} {self.f.ir()})"


class Get(Expr):
    def __init__(self, val: Expr, col: int):
        if val.type.is_struct():
            _check(0 <= col < len(val.type.dtypes), "Get column out of range")
            t = val.type.dtypes[col]
        else:
            _check(val.type.is_primitive() and col == 0, "Get on non-struct")
            t = val.type
        super().__init__(t)
        self.val, self.col = val, col

    def ir(self):
        return f"(get {self.val.ir()} {self.col})"


class New(Expr):
    def __init__(self, vals: Sequence[Expr]):
        vals = list(vals)
        _check(len(vals) > 0, "New with no fields")
        super().__init__(DataType("struct", [v.type for v in vals]))
        self.vals = vals

    def ir(self):
        return "(new " + " ".join(v.ir() for v in self.vals) + ")"


_UNARY = {"not", "abs", "neg", "sqrt", "ceil", "floor"}
_CMP = {"lt", "lte", "gt", "gte", "eq"}


class Nary(Expr):
    def __init__(self, op: str, args: Sequence[Expr]):
        args = list(args)
        a0 = args[0].type
        _check(not a0.is_struct() and not a0.is_vector(), f"{op} on struct/vector")
        if op in _UNARY:
            _check(len(args) == 1, f"{op} is unary")
        else:
            _check(len(args) == 2, f"{op} is binary")
            _check(args[0].type == args[1].type,
                   f"{op} operand types differ: {args[0].type} vs {args[1].type}")
        t = a0
        if op in _CMP:
            t = _bool_t
        elif op in ("not", "and", "or"):
            _check(a0 == _bool_t, f"{op} needs bool operands")
            t = _bool_t
        elif op == "mod":
            _check(not a0.is_float(), "mod of float")
        elif op in ("pow", "ceil", "floor"):
            _check(a0.is_float(), f"{op} of non-float")
        super().__init__(t)
        self.op, self.args = op, args

    def ir(self):
        return f"({self.op} " + " ".join(a.ir() for a in self.args) + ")"


class Element(Expr):
    def __init__(self, vec: Expr, it: Expr):
        _check(vec.type.is_vector(), "Element of non-vector")
        _check(it.type == vec.type.iterty(), "Element iterator type mismatch")
        super().__init__(vec.type.valty())
        self.vec, self.it = vec, it

    def ir(self):
        return f"(elem {self.vec.ir()} {self.it.ir()})"


class In(Expr):
    def __init__(self, it: Expr, vec: Expr):
        _check(it.type == vec.type.iterty(), "In iterator type mismatch")
        super().__init__(_bool_t)
        self.it, self.vec = it, vec

    def ir(self):
        return f"(in {self.it.ir()} {self.vec.ir()})"


class SubVector(Expr):
    def __init__(self, vec: Expr, start: Expr, end: Expr):
        _check(vec.type.is_vector(), "SubVector of non-vector")
        _check(start.type.is_idx() and end.type.is_idx(), "SubVector bounds must be idx")
        super().__init__(vec.type.valty())
        self.vec, self.start, self.end = vec, start, end

    def ir(self):
        return f"(subvec {self.vec.ir()} {self.start.ir()} {self.end.ir()})"


class Length(Expr):
    def __init__(self, vec: Expr, col: int):
        _check(vec.type.is_vector(), "Length of non-vector")
        super().__init__(_idx_t)
        self.vec, self.col = vec, col

    def ir(self):
        return f"(len {self.vec.ir()} {self.col})"


class Locate(Expr):
    """include/reffine/ir/op_to_loop.h:106-117: the row of `it` in vec's index
    (vector_locate, include/reffine/vinstr/vinstr.h:137-146), -1 if absent."""

    def __init__(self, vec: Expr, it: Expr):
        _check(vec.type.is_vector(), "Locate in non-vector")
        _check(vec.type.dim <= 2, "Locate in a vector of more than 2 dimensions")
        _check(it.type == vec.type.iterty(), "Locate iterator type mismatch")
        super().__init__(_idx_t)
        self.vec, self.it = vec, it

    def ir(self):
        return f"(locate {self.vec.ir()} {self.it.ir()})"


class ReadData(Expr):
    """include/reffine/ir/op_to_loop.h:27-41: column `col` (storage index,
    column 0 included) of vec at row `idx`."""

    def __init__(self, vec: Expr, idx: Expr, col: int):
        _check(vec.type.is_vector(), "ReadData of non-vector")
        _check(0 <= col < len(vec.type.dtypes), "ReadData column out of range")
        _check(idx.type.is_idx() or idx.type.is_int(), "ReadData index must be an integer")
        super().__init__(vec.type.dtypes[col])
        self.vec, self.idx, self.col = vec, idx, col

    def ir(self):
        return f"(readdata {self.vec.ir()} {self.idx.ir()} {self.col})"


class InitVal(Expr):
    """include/reffine/ir/expr.h:16-26: `val`, with the symbols `inits`
    computed first."""

    def __init__(self, inits: Sequence[Sym], val: Expr):
        inits = list(inits)
        _check(all(isinstance(i, Sym) for i in inits), "InitVal initialises symbols")
        super().__init__(val.type)
        self.inits, self.val = inits, val

    def ir(self):
        return "(initval (" + " ".join(i.ir() for i in self.inits) + ") " + self.val.ir() + ")"


class BuildIdx(Expr):
    """include/reffine/ir/op_to_loop.h:184-190: `vec` with an index over its
    iterator column (build_vector_index)."""

    def __init__(self, vec: Expr):
        _check(vec.type.is_vector(), "BuildIndex of non-vector")
        super().__init__(vec.type)
        self.vec = vec

    def ir(self):
        return f"(buildidx {self.vec.ir()})"


class Op(Expr):
    """include/reffine/ir/op.h:13-53."""

    def __init__(self, iters: Sequence[Sym], pred: Expr, outputs: Sequence[Expr]):
        iters, outputs = list(iters), list(outputs)
        for it in iters:
            _check(isinstance(it, Sym) and it.type.is_primitive(), "Op iterator")
        for o in outputs:
            _check(o.type.is_val(), "Op output must be a value")
        _check(pred.type == _bool_t, "Op predicate must be bool")
        dt = [i.type for i in iters] + [o.type for o in outputs]
        super().__init__(DataType("vec", dt, len(iters)))
        self.iters, self.pred, self.outputs = iters, pred, outputs

    def ir(self):
        return ("(op (" + " ".join(i.ir() for i in self.iters) + ") " + self.pred.ir() +
                " (" + " ".join(o.ir() for o in self.outputs) + "))")


class Reduce(Expr):
    """include/reffine/ir/op.h:97-114: the lambdas are applied to symbols."""

    def __init__(self, vec: Expr, init: Callable[[], Expr],
                 acc: Callable[[Expr, Expr], Expr]):
        _check(vec.type.is_vector(), "Reduce over non-vector")
        self.vec = vec
        self.init_fn, self.acc_fn = init, acc
        self.init = init()
        super().__init__(self.init.type)
        self.state = Sym("state", self.init.type)
        self.val = Sym("val", vec.type.rowty())
        self.acc = acc(self.state, self.val)
        _check(self.acc.type == self.init.type, "Reduce accumulator changes the state type")

    def ir(self):
        return (f"(reduce {self.vec.ir()} {self.init.ir()} {self.state.ir()} "
                f"{self.val.ir()} {self.acc.ir()})")


class Func:
    """include/reffine/ir/stmt.h:8-27."""

    def __init__(self, name: str, output: Sym, inputs: Sequence[Sym], tbl=None):
        self.name = name
        self.output = output
        self.inputs = list(inputs)
        self.tbl: dict[Sym, Expr] = dict(tbl or {})

    def ir(self) -> str:
        entries = " ".join(f"({s.ir()} {e.ir()})" for s, e in self.tbl.items())
        return (f"(func {self.name} {self.output.ir()} (" +
                " ".join(i.ir() for i in self.inputs) + f") ({entries}))")


# ------------------------------------------------------------ builder names

def _expr(x, like: Expr | None = None) -> Expr:
    return x


def _add(a, b): return Nary("add", [a, b])
def _sub(a, b): return Nary("sub", [a, b])
def _mul(a, b): return Nary("mul", [a, b])
def _div(a, b): return Nary("div", [a, b])
def _max(a, b): return Nary("max", [a, b])
def _min(a, b): return Nary("min", [a, b])
def _mod(a, b): return Nary("mod", [a, b])
def _abs(a): return Nary("abs", [a])
def _neg(a): return Nary("neg", [a])
def _sqrt(a): return Nary("sqrt", [a])
def _pow(a, b): return Nary("pow", [a, b])
def _ceil(a): return Nary("ceil", [a])
def _floor(a): return Nary("floor", [a])
def _lt(a, b): return Nary("lt", [a, b])
def _lte(a, b): return Nary("lte", [a, b])
def _gt(a, b): return Nary("gt", [a, b])
def _gte(a, b): return Nary("gte", [a, b])
def _eq(a, b): return Nary("eq", [a, b])
def _not(a): return Nary("not", [a])
def _and(a, b): return Nary("and", [a, b])
def _or(a, b): return Nary("or", [a, b])


def _const(t: DataType, v) -> Const: return Const(t, v)
def _cast(t: DataType, e: Expr) -> Cast: return Cast(t, e)
def _sel(c, t, f) -> Select: return Select(c, t, f)
def _get(v, k: int) -> Get: return Get(v, k)
def _new(vals: Sequence[Expr]) -> New: return New(vals)
def _elem(vec, it) -> Element: return Element(vec, it)
def _in(it, vec) -> In: return In(it, vec)
def _subvec(vec, start, end) -> SubVector: return SubVector(vec, start, end)
def _len(vec, col: int) -> Length: return Length(vec, col)
def _locate(vec, it) -> Locate: return Locate(vec, it)
def _initval(inits, val) -> InitVal: return InitVal(inits, val)
def _buildidx(vec) -> BuildIdx: return BuildIdx(vec)
def _readdata(vec, idx, col: int) -> ReadData: return ReadData(vec, idx, col)
def _op(iters, pred, outputs) -> Op: return Op(iters, pred, outputs)
def _red(vec, init, acc) -> Reduce: return Reduce(vec, init, acc)
def _sym(name: str, t) -> Sym: return Sym(name, t)
def _func(name: str, output: Sym, inputs: Sequence[Sym]) -> Func: return Func(name, output, inputs)


def _i8(v): return Const(_i8_t, v)
def _i16(v): return Const(_i16_t, v)
def _i32(v): return Const(_i32_t, v)
def _i64(v): return Const(_i64_t, v)
def _u8(v): return Const(_u8_t, v)
def _u16(v): return Const(_u16_t, v)
def _u32(v): return Const(_u32_t, v)
def _u64(v): return Const(_u64_t, v)
def _f32(v): return Const(_f32_t, v)
def _f64(v): return Const(_f64_t, v)
def _ch(v): return Const(_ch_t, v)
def _idx(v): return Const(_idx_t, v)
def _true(): return Const(_bool_t, 1)
def _false(): return Const(_bool_t, 0)


__all__ = [n for n in dir() if n.startswith("_") and not n.startswith("__") and n not in
           ("_ids", "_check", "_PRIM", "_UNARY", "_CMP", "_expr")] + [
    "DataType", "Expr", "Sym", "Func"]
