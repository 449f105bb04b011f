"""The IR text at the boundary: the reference test suite's query functions,
serialised by hand from their C++ builder calls (tests/golden/
ref_queries_ir.json), equal what the Python mirror of the builder writes
(up to symbol ids and table order) and are accepted by the library's
compile check; the reffine-side serialiser (include/reffine_hip_text.hpp)
type-checks against reffine's own IR headers."""
import ctypes
import json
import math
import os
import subprocess

import pytest

from reffine_amd import runtime

import ref_queries as rq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = os.path.join(ROOT, "tests", "golden", "ref_queries_ir.json")
REF_INCLUDE = "/root/reference/include"


def _parse(text):
    """S-expression -> nested lists of atoms."""
    toks = text.replace("(", " ( ").replace(")", " ) ").split()
    pos = 0

    def one():
        nonlocal pos
        t = toks[pos]
        pos += 1
        if t != "(":
            return t
        out = []
        while toks[pos] != ")":
            out.append(one())
        pos += 1
        return out
    e = one()
    assert pos == len(toks)
    return e


def canon(text):
    """A canonical string of a func's IR: table entries ordered by their
    symbol's name, symbol ids renumbered by first appearance, constants as
    floats."""
    f = _parse(text)
    assert f[0] == "func"

    def strip_ids(e):
        if isinstance(e, list):
            if e and e[0] == "sym":
                return ["sym", e[1].split("#")[0], strip_ids(e[2])]
            return [strip_ids(x) for x in e]
        return e
    f[4] = sorted(f[4], key=lambda ent: (ent[0][1].split("#")[0], repr(strip_ids(ent[1]))))
    ids = {}

    def walk(e):
        if isinstance(e, list):
            if e and e[0] == "sym":
                name, old = e[1].split("#")
                if e[1] not in ids:
                    ids[e[1]] = f"{name}#{len(ids) + 1}"
                return f"(sym {ids[e[1]]} {walk(e[2])})"
            if e and e[0] == "const":
                v = e[2]
                fv = float(v) if v not in ("inf", "-inf", "nan") else v
                return f"(const {walk(e[1])} {fv!r})" if isinstance(fv, float) and not math.isnan(fv) \
                    else f"(const {walk(e[1])} {v})"
            return "(" + " ".join(walk(x) for x in e) + ")"
        return e
    return walk(f)


def fixtures():
    with open(FIXTURES) as fh:
        d = json.load(fh)
    return {k: v for k, v in d.items() if not k.startswith("_")}


BUILDERS = {"vector_op": rq.vector_op, "transform_op": rq.transform_op, "join_op": rq.join_op,
            "nested_op": rq.nested_op, "multidim_op": rq.multidim_op,
            "gen_table_op": rq.gen_table_op}


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_mirror_writes_the_reference_queries(name):
    want = fixtures()[name]["ir"]
    assert canon(BUILDERS[name]().ir()) == canon(want)


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_fixture_text_compiles_for_gfx950(name):
    buf = ctypes.create_string_buffer(1 << 20)
    st = runtime.lib().rf_compile_check(fixtures()[name]["ir"].encode(), 0, buf, len(buf))
    assert st == 0, runtime.lib().rf_last_error()


def test_canon_tells_queries_apart():
    """The comparison is not vacuous: a changed constant, a swapped operand
    or a different column differ after canonicalisation."""
    base = canon(fixtures()["transform_op"]["ir"])
    for a, b in (("(const i64 10)", "(const i64 11)"), ("(get (sym elem#3", "(get (sym elem#3"),
                 ("(gt (sym t#1 i64) (const i64 5))", "(gt (const i64 5) (sym t#1 i64))")):
        txt = fixtures()["transform_op"]["ir"].replace(a, b)
        if a != b:
            assert canon(txt) != base
    swapped = fixtures()["gen_table_op"]["ir"].replace("(lt (sym t#1 i64) (sym ub#3 i64))",
                                                        "(lt (sym t#1 i64) (sym lb#2 i64))")
    assert canon(swapped) != canon(fixtures()["gen_table_op"]["ir"])


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reffine's headers are not here")
def test_serializer_type_checks_against_reffine_headers():
    """include/reffine_hip_text.hpp and a reffine program that uses it
    (tests/c_abi/rf_text_example.cpp) compile against reffine's IR, builder
    and visitor headers (syntax and types; linking needs a reffine build)."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
                        "-I", REF_INCLUDE, os.path.join(ROOT, "tests", "c_abi",
                                                        "rf_text_example.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
