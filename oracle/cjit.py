"""CPU JIT restatement of reffine's query compiler (test infrastructure).

Follows the reference's CPU path for the scan idioms:
  * LoopGen (src/pass/loopgen.cpp:206-269, 113-204, 44-111): a Reduce over a
    SubVector loops i in [start, end) with body condition = validity of
    column 1 and row = columns 1..; a Reduce / Op over `_in(t, T)` loops
    over T's storage index with t = column 0 and body condition = validity
    of column 0 (VecSpace, src/iter/iter_space.cpp:55-91); further `_in`
    terms are intersections (InterSpace), here evaluated as key lookups;
    `vectorize=True` folds `s = cond ? acc : s`, otherwise `if (cond) s =
    acc`; an Op compacts qualifying rows at out_idx.
  * CanonPass / ScalarPass (src/pass/canonpass.cpp:31-64,
    src/pass/scalarpass.cpp:103-227): state fields become scalars, the
    output struct is written field by field in declaration order.
  * CEmitter (src/pass/cemitter.cpp:80-252): infix C text, casts
    `(T) x`, constants printed with std::to_string (so a double constant
    keeps 6 decimals), `#pragma clang loop vectorize(enable)
    interleave(enable)` on a `while (1)` loop.
  * LLVMGen::parse (src/pass/llvmgen.cpp:653-696): `clang++ -O3` on the text
    (generic x86-64, no -march, hence no FMA); here into a shared object
    loaded with ctypes instead of the ORC JIT.
Predicates over non-iterator columns inside an Op are evaluated per row
(the GPU path's semantics; the reference's Reffine pass only turns iterator
bounds into loop bounds).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

from reffine_amd import reffiner as R

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
CXX = os.environ.get("RF_ORACLE_CXX", "/opt/rocm/llvm/bin/clang++")

_CTYPE = {"b": "bool", "i8": "int8_t", "u8": "uint8_t", "i16": "int16_t", "u16": "uint16_t",
          "i32": "int32_t", "u32": "uint32_t", "i64": "int64_t", "u64": "uint64_t",
          "f32": "float", "f64": "double", "x": "int64_t"}
_NP = {"b": np.int8, "i8": np.int8, "u8": np.uint8, "i16": np.int16, "u16": np.uint16,
       "i32": np.int32, "u32": np.uint32, "i64": np.int64, "u64": np.uint64,
       "f32": np.float32, "f64": np.float64, "x": np.int64}

PRELUDE = r"""
#include <stdint.h>
#include <math.h>
#include <algorithm>
#include <cmath>
#include <limits>
struct RfTab { const void* const* data; const uint16_t* const* valid; int64_t n; };
struct RfIdx { const int64_t* keys; const int64_t* rows; uint64_t mask; int64_t min_row; };
static inline bool get_null_bit(const uint16_t* bm, int64_t i) {
    return !bm || (bm[i >> 4] & (1u << (i & 15)));
}
static inline uint64_t rf_hash(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33; return k;
}
static inline int64_t vector_locate(const RfIdx* h, int64_t key) {
    if (key == (int64_t)0x8000000000000000ull) return h->min_row;
    uint64_t s = rf_hash((uint64_t)key) & h->mask;
    for (;;) {
        int64_t k = h->keys[s];
        if (k == key) return h->rows[s];
        if (k == (int64_t)0x8000000000000000ull) return -1;
        s = (s + 1) & h->mask;
    }
}
"""


def cpp_const(t: R.DataType, v: float) -> str:
    """CEmitter::visit(Const) (src/pass/cemitter.cpp:80-103)."""
    ty = _CTYPE[t.btype]
    if t.btype == "b":
        return "true" if v else "false"
    if t.is_float():
        return f"{ty}({v:f})"
    if t.btype.startswith("u"):
        return f"{ty}({int(v) & 0xFFFFFFFFFFFFFFFF})"
    return f"{ty}({int(v)})"


_FIXED_IDENT = {
    ("add", True): "-0.0", ("add", False): "0",
    ("max", True): "-INFINITY", ("min", True): "INFINITY",
}


def _ident(op: str, t: R.DataType) -> str:
    """The identity a fixed-order fold starts each lane and thread from
    (-0.0 for float sums: the exact additive identity)."""
    ct = _CTYPE[t.btype]
    if (op, t.is_float()) in _FIXED_IDENT:
        return f"({ct})({_FIXED_IDENT[(op, t.is_float())]})"
    lim = "min" if op == "max" else "max"
    return f"std::numeric_limits<{ct}>::{lim}()"


def _combine(op: str, a: str, b: str) -> str:
    """a (+) b with std::max / std::min semantics for max / min."""
    if op == "add":
        return f"({a} + {b})"
    if op == "max":
        return f"(({a} < {b}) ? {b} : {a})"
    return f"(({b} < {a}) ? {b} : {a})"


class _Gen:
    def __init__(self, fn: R.Func, vectorize: bool, fixed=None):
        self.fn = fn
        self.vectorize = vectorize
        # (BLOCK, ITEMS, SUB, ops per state field or None for all "add"):
        # emit the reductions in the GPU kernels' deterministic order.
        self.fixed = fixed
        self.tbl = {s: e for s, e in fn.tbl.items()}
        self.inputs = list(fn.inputs)
        self.lines: list[str] = []
        self.nt = 0
        self.probes: list[tuple[int, str]] = []  # (table input, key text)

    # ------------------------------------------------------------ helpers
    def inp(self, s) -> int:
        for i, x in enumerate(self.inputs):
            if x is s:
                return i
        return -1

    def resolve(self, e):
        while isinstance(e, R.Sym) and self.inp(e) < 0 and e in self.tbl:
            e = self.tbl[e]
        return e

    def tmp(self, rhs: str) -> str:
        v = f"t{self.nt}"
        self.nt += 1
        self.lines.append(f"        const auto {v} = {rhs};")
        return v

    # ------------------------------------------------------------ values
    def col(self, ti: int, c: int, idx: str) -> str:
        t = self.inputs[ti].type.dtypes[c]
        return f"((const {_CTYPE[t.btype]}*){self.tname(ti)}->data[{c}])[{idx}]"

    def tname(self, ti: int) -> str:
        return f"in{ti}"

    def val(self, e, env) -> object:
        """Returns C text for primitives, or a list for struct values."""
        if isinstance(e, R.Sym):
            if e in env:
                return env[e]
            i = self.inp(e)
            if i >= 0:
                return f"in{i}"
            if e in self.tbl:
                return self.val(self.tbl[e], env)
            raise KeyError(e.name)
        if isinstance(e, R.Const):
            return cpp_const(e.type, e.val)
        if isinstance(e, R.Cast):
            return self.tmp(f"({_CTYPE[e.type.btype]}) {self.val(e.arg, env)}")
        if isinstance(e, R.Get):
            v = self.val(e.val, env)
            return v[e.col] if isinstance(v, list) else v
        if isinstance(e, R.New):
            return [self.val(x, env) for x in e.vals]
        if isinstance(e, R.Select):
            c = self.val(e.cond, env)
            t, f = self.val(e.t, env), self.val(e.f, env)
            if isinstance(t, list):
                return [self.tmp(f"({c} ? {a} : {b})") for a, b in zip(t, f)]
            return self.tmp(f"({c} ? {t} : {f})")
        if isinstance(e, R.Nary):
            a = [self.val(x, env) for x in e.args]
            ty = _CTYPE[e.type.btype]
            ops = {"add": "+", "sub": "-", "mul": "*", "div": "/", "mod": "%", "lt": "<",
                   "lte": "<=", "gt": ">", "gte": ">=", "eq": "==", "and": "&&", "or": "||"}
            if e.op in ops:
                return self.tmp(f"({a[0]} {ops[e.op]} {a[1]})")
            if e.op == "not":
                return self.tmp(f"!{a[0]}")
            if e.op == "neg":
                return self.tmp(f"-{a[0]}")
            if e.op in ("max", "min"):
                return self.tmp(f"std::{e.op}<{ty}>({a[0]}, {a[1]})")
            if e.op == "abs":
                return self.tmp(f"(({a[0]}) < 0 ? -({a[0]}) : ({a[0]}))")
            if e.op in ("sqrt", "ceil", "floor"):
                return self.tmp(f"std::{e.op}({a[0]})")
            if e.op == "pow":
                return self.tmp(f"std::pow({a[0]}, {a[1]})")
            raise NotImplementedError(e.op)
        if isinstance(e, R.Length):
            i = self.inp(self.resolve(e.vec))
            return f"in{i}->n"
        if isinstance(e, R.Element):
            vec = self.resolve(e.vec)
            ti = self.inp(vec)
            itv = self.val(e.it, env)
            drv = env.get("__driver")
            if drv is not None and ti == drv[0] and isinstance(e.it, R.Sym) and e.it is drv[1]:
                idx = "i"
            else:
                idx = self.probe(ti, itv)
            return [self.col(ti, c, idx) if idx == "i" else
                    self.tmp(f"({idx} >= 0 ? {self.col(ti, c, idx)} : "
                             f"({_CTYPE[vec.type.dtypes[c].btype]})0)")
                    for c in range(vec.type.dim, len(vec.type.dtypes))]
        if isinstance(e, R.In):
            vec = self.resolve(e.vec)
            ti = self.inp(vec)
            drv = env.get("__driver")
            if drv is not None and ti == drv[0] and isinstance(e.it, R.Sym) and e.it is drv[1]:
                return "true"
            h = self.probe(ti, self.val(e.it, env))
            return f"({h} >= 0)"
        raise NotImplementedError(type(e).__name__)

    def probe(self, ti: int, key: str) -> str:
        for j, (t, k) in enumerate(self.probes):
            if t == ti and k == key:
                return f"h{j}"
        j = len(self.probes)
        self.probes.append((ti, key))
        self.lines.append(f"        const int64_t h{j} = vector_locate(idx{ti}, (int64_t){key});")
        return f"h{j}"

    # ------------------------------------------------------------- loops
    def source(self, vec):
        """(loop header lines, row env builder, body condition builder)."""
        vec = self.resolve(vec)
        if isinstance(vec, R.SubVector):
            T = self.resolve(vec.vec)
            ti = self.inp(T)
            start = self.val(vec.start, {})
            end = self.val(vec.end, {})

            def row(env):
                return [self.col(ti, c, "i") for c in range(1, len(T.type.dtypes))]

            def cond(env):
                return f"get_null_bit(in{ti}->valid[1], i)"
            return ti, start, end, row, cond, None
        if isinstance(vec, R.Op):
            it = vec.iters[0]
            conj = []

            def flat(p):
                p = self.resolve(p)
                if isinstance(p, R.Nary) and p.op == "and":
                    flat(p.args[0])
                    flat(p.args[1])
                else:
                    conj.append(p)
            flat(vec.pred)
            ti = -1
            for c in conj:
                if isinstance(c, R.In) and c.it is it and self.inp(self.resolve(c.vec)) >= 0:
                    ti = self.inp(self.resolve(c.vec))
                    break
            if ti < 0:
                return self.range_source(vec, it, conj)
            T = self.inputs[ti]

            def row(env):
                env[it] = self.col(ti, 0, "i")
                env["__driver"] = (ti, it)
                return [self.val(x, env) for x in [it] + vec.outputs]

            def cond(env):
                env[it] = self.col(ti, 0, "i")
                env["__driver"] = (ti, it)
                parts = [f"get_null_bit(in{ti}->valid[0], i)"]
                for c in conj:
                    parts.append(self.val(c, env))
                return " && ".join(f"({p})" for p in parts)
            return ti, "int64_t(0)", f"in{ti}->n", row, cond, vec
        raise NotImplementedError(f"cjit: reduce over {type(vec).__name__}")

    def range_source(self, op, it, conj):
        """Universal space bounded by lb <= t < ub (LBound/UBound spaces,
        src/iter/iter_space.cpp:150-200): i runs over [lb, ub), t = i."""
        lb = ub = None
        rest = []
        ct = _CTYPE[it.type.btype]
        for c in conj:
            if isinstance(c, R.Nary) and c.op in ("gte", "gt", "lt", "lte"):
                a0, a1 = c.args
                if a0 is it and a1 is not it:
                    if c.op == "gte" and lb is None:
                        lb = self.val(a1, {})
                        continue
                    if c.op == "gt" and lb is None:
                        lb = f"({self.val(a1, {})} + {ct}(1))"
                        continue
                    if c.op == "lt" and ub is None:
                        ub = self.val(a1, {})
                        continue
                    if c.op == "lte" and ub is None:
                        ub = f"({self.val(a1, {})} + {ct}(1))"
                        continue
            rest.append(c)
        if lb is None or ub is None:
            raise NotImplementedError("cjit: unbounded Op iterator")

        def row(env):
            env[it] = f"(({ct}) i)"
            return [self.val(x, env) for x in [it] + op.outputs]

        def cond(env):
            env[it] = f"(({ct}) i)"
            parts = ["true"] + [self.val(c, env) for c in rest]
            return " && ".join(f"({p})" for p in parts)
        return -1, f"(int64_t){lb}", f"(int64_t){ub}", row, cond, op

    def state_fields(self, t: R.DataType, prefix: list):
        if t.is_struct():
            out = []
            for i, d in enumerate(t.dtypes):
                out += self.state_fields(d, prefix + [i])
            return out
        return [(prefix, t)]

    def nest(self, t, flat_vals, k=None):
        if k is None:
            k = [0]
        if t.is_struct():
            return [self.nest(d, flat_vals, k) for d in t.dtypes]
        v = flat_vals[k[0]]
        k[0] += 1
        return v

    def flatten(self, v):
        if isinstance(v, list):
            out = []
            for x in v:
                out += self.flatten(x)
            return out
        return [v]

    def components(self):
        comps = []

        def walk(e):
            e2 = self.resolve(e) if isinstance(e, R.Sym) else e
            if isinstance(e2, (R.Reduce, R.Op)):
                if all(c is not e2 for c in comps):
                    comps.append(e2)
            elif isinstance(e2, R.New):
                for x in e2.vals:
                    walk(x)
            elif isinstance(e2, R.Get):
                walk(e2.val)
            elif isinstance(e2, R.Cast):
                walk(e2.arg)
            elif isinstance(e2, R.Nary):
                for x in e2.args:
                    walk(x)
        walk(self.fn.output)
        return comps

    def gen(self) -> tuple[str, dict]:
        comps = self.components()
        params = ["void* output_addr"]
        for i, s in enumerate(self.inputs):
            if s.type.is_vector():
                params += [f"const RfTab* in{i}", f"const RfIdx* idx{i}"]
            else:
                params.append(f"{_CTYPE[s.type.btype]} in{i}")
        blocks = []
        meta = {"comps": []}
        for ci, comp in enumerate(comps):
            self.lines = []
            self.probes = []
            vec = comp.vec if isinstance(comp, R.Reduce) else comp
            ti, start, end, row, cond, op = self.source(vec)
            if isinstance(comp, R.Reduce):
                fields = self.state_fields(comp.type, [])
                init_vals = self.flatten(self.val(comp.init, {}))
            head, self.lines = self.lines, []
            env = {}
            c = cond(env)
            vals = row(env)
            blk = list(head)
            if isinstance(comp, R.Reduce):
                svars = [f"s_{k}" for k in range(len(fields))]
                env[comp.state] = self.nest(comp.type, svars)
                env[comp.val] = vals
                accv = self.flatten(self.val(comp.acc, env))
                if self.fixed is not None:
                    blk += self.fixed_fold(fields, svars, init_vals, accv, c, start, end)
                    off = 0
                    for (path, t), sv in zip(fields, svars):
                        w = np.dtype(_NP[t.btype]).itemsize
                        off = (off + w - 1) // w * w
                        blk.append(f"    *({_CTYPE[t.btype]}*)((char*)state{ci} + {off}) = {sv};")
                        off += w
                    meta["comps"].append(("reduce", comp, fields))
                    blocks.append(blk)
                    meta.setdefault("probed", set()).update(t for t, _ in self.probes)
                    continue
                for (path, t), sv, iv in zip(fields, svars, init_vals):
                    blk.append(f"    {_CTYPE[t.btype]} {sv} = {iv};")
                blk += [f"    int64_t i = {start};", f"    const int64_t end = {end};",
                        "    #pragma clang loop vectorize(enable) interleave(enable)",
                        "    while (1) {", "        if (i >= end) break;"]
                blk += self.lines
                if self.vectorize:  # loopgen.cpp:253-259
                    blk += [f"        {sv} = ({c}) ? {a} : {sv};" for sv, a in zip(svars, accv)]
                else:  # loopgen.cpp:260-262 + canonpass.cpp:42-44
                    blk += [f"        if ({c}) {{"]
                    blk += [f"            {sv} = {a};" for sv, a in zip(svars, accv)]
                    blk += ["        }"]
                blk += ["        i = i + int64_t(1);", "    }"]
                off = 0
                for (path, t), sv in zip(fields, svars):
                    w = np.dtype(_NP[t.btype]).itemsize
                    off = (off + w - 1) // w * w
                    blk.append(f"    *({_CTYPE[t.btype]}*)((char*)state{ci} + {off}) = {sv};")
                    off += w
                meta["comps"].append(("reduce", comp, fields))
            else:
                types = [x.type for x in [op.iters[0]] + op.outputs]
                blk += ["    int64_t out_idx = 0;", f"    int64_t i = {start};",
                        f"    const int64_t end = {end};",
                        "    #pragma clang loop vectorize(enable) interleave(enable)",
                        "    while (1) {", "        if (i >= end) break;"]
                blk += self.lines + [f"        if ({c}) {{"]
                for k, (v, t) in enumerate(zip(vals, types)):
                    ct = _CTYPE[t.btype]
                    blk.append(f"            (({ct}*)outcols{ci}[{k}])[out_idx] = ({ct}){v};")
                blk += ["            out_idx = out_idx + 1;", "        }",
                        "        i = i + int64_t(1);", "    }", f"    *outcount{ci} = out_idx;"]
                meta["comps"].append(("op", comp, types))
            blocks.append(blk)
            meta.setdefault("probed", set()).update(t for t, _ in self.probes)
        for ci, comp in enumerate(comps):
            if isinstance(comp, R.Op):
                params += [f"void* const* outcols{ci}", f"int64_t* outcount{ci}"]
        for ci, comp in enumerate(comps):
            if isinstance(comp, R.Reduce):
                params.append(f"void* state{ci}")
        text = PRELUDE + f'\nextern "C" void {self.fn.name}({", ".join(params)}) {{\n'
        for blk in blocks:
            text += "  {\n" + "\n".join(blk) + "\n  }\n"
        text += "}\n"
        return text, meta


def _fixed_fold(self, fields, svars, init_vals, accv, c, start, end):
    """C text of a reduction in the GPU scan's deterministic order: rows
    [base, end) in wave-tiles of SUB*64*ITEMS rows (base = start rounded
    down to a multiple of ITEMS); in sub-step s lane l folds rows
    tile0 + (s*64 + l)*ITEMS + [0, ITEMS) in ascending order from the
    identity, applying the query's own accumulator to its lane state; the
    lanes combine by xor butterflies (lane_i = lane_i (+) lane_{i^o},
    o = 1, 2, ..., 32) and lane 0 is the tile's partial; thread t of a block of
    BLOCK threads folds partials t, t + BLOCK, ... from the identity; each
    wave combines its threads by the same butterflies; the waves fold in
    order; the result is init (+) total."""
    block, items, sub, ops = self.fixed
    ops = ops or ["add"] * len(fields)
    if len(ops) != len(fields):
        raise ValueError("one combine operator per state field")
    L = []
    ty = [_CTYPE[t.btype] for _, t in fields]
    idn = [_ident(op, t) for op, (_, t) in zip(ops, fields)]
    L += [f"    const int64_t lo = {start};", f"    const int64_t end = {end};",
          f"    const int64_t TILE = (int64_t){sub} * 64 * {items};",
          f"    const int64_t base = lo >= 0 ? lo - lo % {items} : "
          f"-(((-lo) + {items} - 1) / {items}) * {items};",
          "    const int64_t ntiles = end > lo ? (end - base + TILE - 1) / TILE : 0;"]
    for k, (t_, idk) in enumerate(zip(ty, idn)):
        L.append(f"    {t_}* part{k} = new {t_}[ntiles > 0 ? ntiles : 1];")
    L.append("    for (int64_t tile = 0; tile < ntiles; tile++) {")
    for k, (t_, idk) in enumerate(zip(ty, idn)):
        L.append(f"        {t_} ln{k}[64];")
        L.append(f"        for (int l = 0; l < 64; l++) ln{k}[l] = {idk};")
    L += [f"        for (int s = 0; s < {sub}; s++) {{",
          "        for (int l = 0; l < 64; l++) {"]
    for k, (sv, t_) in enumerate(zip(svars, ty)):
        L.append(f"            {t_} {sv} = ln{k}[l];")
    L += [f"            for (int k = 0; k < {items}; k++) {{",
          f"                const int64_t i = base + tile * TILE + ((int64_t)s * 64 + l) * {items} + k;",
          "                if (i < lo || i >= end) continue;"]
    L += ["        " + x for x in self.lines]
    L.append(f"                if ({c}) {{")
    L += [f"                    const {t_} n{k} = {a};" for k, (a, t_) in enumerate(zip(accv, ty))]
    L += [f"                    {sv} = n{k};" for k, sv in enumerate(svars)]
    L += ["                }", "            }"]
    for k, sv in enumerate(svars):
        L.append(f"            ln{k}[l] = {sv};")
    L += ["        }", "        }"]
    for k, (t_, op) in enumerate(zip(ty, ops)):
        L += ["        for (int o = 1; o < 64; o <<= 1) {",
              f"            {t_} nx[64];",
              f"            for (int l = 0; l < 64; l++) nx[l] = {_combine(op, f'ln{k}[l]', f'ln{k}[l ^ o]')};",
              f"            for (int l = 0; l < 64; l++) ln{k}[l] = nx[l];",
              "        }",
              f"        part{k}[tile] = ln{k}[0];"]
    L.append("    }")
    for k, (sv, t_, op, idk, iv) in enumerate(zip(svars, ty, ops, idn, init_vals)):
        L += ["    {",
              f"        {t_} th[{block}];",
              f"        for (int t = 0; t < {block}; t++) th[t] = {idk};",
              f"        for (int64_t p = 0; p < ntiles; p++) th[p % {block}] = "
              f"{_combine(op, f'th[p % {block}]', f'part{k}[p]')};",
              f"        for (int w = 0; w < {block} / 64; w++) {{",
              "            for (int o = 1; o < 64; o <<= 1) {",
              f"                {t_} nx[64];",
              f"                for (int l = 0; l < 64; l++) nx[l] = "
              f"{_combine(op, 'th[w * 64 + l]', 'th[w * 64 + (l ^ o)]')};",
              "                for (int l = 0; l < 64; l++) th[w * 64 + l] = nx[l];",
              "            }",
              "        }",
              f"        {t_} tot = th[0];",
              f"        for (int w = 1; w < {block} / 64; w++) tot = {_combine(op, 'tot', 'th[w * 64]')};",
              f"        {t_} {sv} = ({t_})({iv});",
              f"        {sv} = {_combine(op, sv, 'tot')};",
              f"        part{k}[0] = {sv};",
              "    }"]
    # the final states, declared at block scope for the state write
    for k, (sv, t_) in enumerate(zip(svars, ty)):
        L.append(f"    const {t_} {sv} = part{k}[0];")
    for k in range(len(fields)):
        L.append(f"    delete[] part{k};")
    return L


_Gen.fixed_fold = _fixed_fold


class CompiledQuery:
    """A query compiled by the restated CPU JIT. Call with numpy tables."""

    def __init__(self, fn: R.Func, vectorize: bool = False, fixed_order=None):
        self.fn = fn
        gen = _Gen(fn, vectorize, fixed_order)
        self.src, self.meta = gen.gen()
        os.makedirs(BUILD, exist_ok=True)
        h = hashlib.sha1(self.src.encode()).hexdigest()[:16]
        so = os.path.join(BUILD, f"cjit_{fn.name}_{h}.so")
        if not os.path.exists(so):
            cpp = so[:-3] + ".cpp"
            with open(cpp, "w") as f:
                f.write(self.src)
            r = subprocess.run([CXX, "-O3", "-shared", "-fPIC", "-std=c++17", cpp, "-o", so + ".tmp"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("cjit compile failed:\n" + r.stderr + "\n" + self.src)
            os.replace(so + ".tmp", so)
        self.lib = ctypes.CDLL(so)
        self.cfn = getattr(self.lib, fn.name)
        self.cfn.restype = None

    def __call__(self, *inputs, cap: int | None = None):
        """inputs: per func input, either a dict {"cols": [np arrays],
        "valid": [bool arrays|None], "index": (keys, rows, mask, min_row)}
        or a scalar. Returns (reduce states..., selections...) per component."""
        keep = []
        args = [ctypes.c_void_p(0)]  # output_addr (unused; states are written per component)
        for ti, (s, x) in enumerate(zip(self.fn.inputs, inputs)):
            if s.type.is_vector():
                cols = [np.ascontiguousarray(c) for c in x["cols"]]
                keep += cols
                data = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
                valid = (ctypes.c_void_p * len(cols))()
                for k, v in enumerate(x.get("valid") or [None] * len(cols)):
                    if v is not None:
                        w = _bitmap_u16(v)
                        keep.append(w)
                        valid[k] = w.ctypes.data
                tab_s = _RfTab(ctypes.cast(data, ctypes.c_void_p), ctypes.cast(valid, ctypes.c_void_p),
                               len(cols[0]))
                keep += [data, valid, tab_s]
                args.append(ctypes.byref(tab_s))
                idx = x.get("index")
                probed = ti in self.meta.get("probed", set())
                if idx is None and probed and s.type.dtypes[0].btype == "i64":
                    idx = build_index(cols[0], (x.get("valid") or [None])[0])
                if idx is not None:
                    keys, rows, mask, mn = idx
                    keep += [keys, rows]
                    ix = _RfIdx(keys.ctypes.data, rows.ctypes.data, mask, mn)
                    keep.append(ix)
                    args.append(ctypes.byref(ix))
                else:
                    args.append(ctypes.c_void_p(0))
            elif s.type.is_float():
                args.append(ctypes.c_double(float(x)))
            else:
                args.append(ctypes.c_int64(int(x)))
        results = []
        states = []
        for kind, comp, info in self.meta["comps"]:
            if kind == "op":
                sizes = [len(i["cols"][0]) for i in inputs if isinstance(i, dict)]
                n = cap if cap is not None else max(sizes + [1])
                bufs = [np.empty(max(n, 1), dtype=_NP[t.btype]) for t in info]
                keep += bufs
                arr = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
                cnt = np.zeros(1, dtype=np.int64)
                keep += [arr, cnt]
                args.append(arr)
                args.append(ctypes.c_void_p(cnt.ctypes.data))
                results.append(("op", bufs, cnt))
        for kind, comp, info in self.meta["comps"]:
            if kind == "reduce":
                buf = ctypes.create_string_buffer(8 * len(info) + 64)
                keep.append(buf)
                args.append(buf)
                states.append((comp, info, buf))
        self.cfn(*args)
        out = []
        si = 0
        ri = 0
        for kind, comp, info in self.meta["comps"]:
            if kind == "reduce":
                c, fields, buf = states[si]
                si += 1
                vals = []
                off = 0
                for path, t in fields:
                    w = np.dtype(_NP[t.btype]).itemsize
                    off = (off + w - 1) // w * w
                    vals.append(np.frombuffer(buf.raw, dtype=_NP[t.btype], count=1,
                                              offset=off)[0].item())
                    off += w
                out.append(tuple(vals))
            else:
                _, bufs, cnt = results[ri]
                ri += 1
                k = int(cnt[0])
                out.append([b[:k].copy() for b in bufs])
        return out


class _RfTab(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("valid", ctypes.c_void_p), ("n", ctypes.c_int64)]


class _RfIdx(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_void_p), ("rows", ctypes.c_void_p), ("mask", ctypes.c_uint64),
                ("min_row", ctypes.c_int64)]


def _bitmap_u16(valid: np.ndarray) -> np.ndarray:
    b = np.packbits(np.asarray(valid, dtype=bool), bitorder="little")
    if len(b) % 2:
        b = np.concatenate([b, np.zeros(1, np.uint8)])
    return np.concatenate([b, np.zeros(2, np.uint8)])


def build_index(keys: np.ndarray, valid=None):
    """ArrowTable2::build_index restated (table.h:107-138): key -> first valid
    row, in an open-addressing table with the kernels' hash."""
    keys = np.asarray(keys, dtype=np.int64)
    n = len(keys)
    cap = 16
    while cap < 2 * n:
        cap <<= 1
    empty = np.int64(-0x8000000000000000)
    tk = np.full(cap, empty, dtype=np.int64)
    tr = np.full(cap, np.iinfo(np.int64).max, dtype=np.int64)
    ok = np.ones(n, bool) if valid is None else np.asarray(valid, bool)
    rows = np.nonzero(ok)[0]
    kk = keys[rows]
    # first occurrence per key
    uk, first = np.unique(kk, return_index=True)
    urow = rows[first]
    mn = -1
    m = uk == empty
    if m.any():
        mn = int(urow[m][0])
        uk, urow = uk[~m], urow[~m]
    with np.errstate(over="ignore"):
        h = uk.view(np.uint64).copy()
        h ^= h >> np.uint64(33)
        h *= np.uint64(0xff51afd7ed558ccd)
        h ^= h >> np.uint64(33)
        h *= np.uint64(0xc4ceb9fe1a85ec53)
        h ^= h >> np.uint64(33)
    slot = (h & np.uint64(cap - 1)).astype(np.int64)
    for k, r, s in zip(uk.tolist(), urow.tolist(), slot.tolist()):
        while tk[s] != empty:
            s = (s + 1) & (cap - 1)
        tk[s] = k
        tr[s] = r
    return tk, tr, cap - 1, mn


def compile_op(fn: R.Func, vectorize: bool = False, fixed_order=None) -> CompiledQuery:
    """fixed_order=(BLOCK, ITEMS, SUB[, ops]): the reductions in the GPU
    deterministic mode's order (ops: each state field's combine, "add" by
    default)."""
    if fixed_order is not None and len(fixed_order) == 3:
        fixed_order = tuple(fixed_order) + (None,)
    return CompiledQuery(fn, vectorize, fixed_order)


def build() -> None:
    """Checks the oracle's compiler works (compiles a trivial query)."""
    os.makedirs(BUILD, exist_ok=True)


def c2_query(d0=1000, vlo=0.05, vhi=0.07, qmax=24):
    """The scan-select workload compiled by the restated CPU JIT; returns a
    callable (rowid, v, d, q) -> (count, sum_q, sum_vq, selection)."""
    from reffine_amd import workloads
    fn = workloads.scan_select_c2(d0, vlo, vhi, qmax)
    cq = CompiledQuery(fn)

    def run(rowid, v, d, q):
        k = np.zeros(len(v), np.int64)
        st, sel = cq({"cols": [rowid, k, v, d, q]})
        return st[0], st[1], st[2], sel[0]
    return run
