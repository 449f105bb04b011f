"""C ABI checks that need no GPU: the library loads, exports exactly what
include/reffine_hip.h declares, rejects bad input with status codes, and
every workload query lowers and compiles for gfx950."""
import ctypes
import os
import re

import pytest

from reffine_amd import runtime
from reffine_amd.reffiner import (_add, _f64, _f64_t, _func, _get, _i64, _i64_t, _idx, _len,
                                  _mul, _new, _red, _sel, _subvec, _sym, _vec_t)
from reffine_amd import workloads

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "reffine_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(rf_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = runtime.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in reffine_hip.h but not exported"
    assert sorted(runtime.EXPORTED) == names


def test_errors_are_status_codes_without_device():
    lib = runtime.lib()
    h = ctypes.c_void_p()
    # No rf_init: compile needs a device -> HIP error status, no crash.
    st = lib.rf_compile(workloads.scan_agg_c1().ir().encode(), 0, ctypes.byref(h))
    assert st != 0
    assert b"rf_init" in lib.rf_last_error()
    # Malformed IR is RF_ERR_INVALID.
    buf = ctypes.create_string_buffer(256)
    assert lib.rf_compile_check(b"(func", 0, buf, 256) == 1
    assert lib.rf_compile_check(b"(func f (sym o#1 i64) () ())", 0, buf, 256) == 2


def test_serve_checks_its_settings_before_any_device_work():
    """rf_serve refuses a runner count outside 2..4 (0 picks the default)
    with RF_ERR_INVALID, before it needs a device."""
    lib = runtime.lib()
    wall = ctypes.c_double()
    for runners in (1, 5, -2):
        assert lib.rf_serve(None, 0, 16, 1 << 20, 0, runners, ctypes.byref(wall)) == 1
        assert b"runners must be 2..4" in lib.rf_last_error()


def test_non_associative_fold_lowers_to_a_sequential_fold():
    """s * 0.5 + v is not a fold the kernels can re-associate: it lowers to
    the sequential fold (one wave, rows in order: the reference's loop),
    which runs as a query of its own beside other components."""
    T = _sym("T", _vec_t(2, _i64_t, _i64_t, _f64_t))
    red = _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _f64(0),
               lambda s, v: _add(_mul(s, _f64(0.5)), _get(v, 1)))
    rs = _sym("r", red)
    fn = _func("seq", rs, [T])
    fn.tbl[rs] = red
    text = runtime.compile_check(fn)
    assert "sequential fold" in text and "rf_qs" in text
    # Beside an associative fold in one output: one query per component.
    red2 = _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _i64(0),
                lambda s, v: _add(s, _get(v, 0)))
    rs2 = _sym("r2", red2)
    both = _new([rs, rs2])
    bs = _sym("both", both)
    fn2 = _func("mixed", bs, [T])
    fn2.tbl[rs] = red
    fn2.tbl[rs2] = red2
    fn2.tbl[bs] = both
    text = runtime.compile_check(fn2)
    assert "independent components" in text and "rf_qs" in text


@pytest.mark.parametrize("det", [False, True])
@pytest.mark.parametrize("build", [workloads.scan_agg_c1, workloads.scan_select_c2,
                                   lambda: workloads.scan_select_c2(with_selection=False)])
def test_workloads_compile_for_gfx950(build, det):
    s = runtime.compile_check(build(), deterministic=det)
    assert "code object bytes" in s


def test_nullable_variant_compiles():
    s = runtime.compile_check(workloads.scan_agg_c1(), nullable=True)
    assert "rf_valid_bits" in s


def test_reference_test_queries_compile_for_gfx950():
    import ref_queries as rq
    for build in [rq.sum_hours_subvec, rq.vector_op, rq.transform_op, rq.gen_table_op,
                  rq.join_op]:
        assert "code object bytes" in runtime.compile_check(build())


def test_range_bounds_are_extracted():
    import ref_queries as rq
    s = runtime.compile_check(rq.gen_table_op())
    assert "scan range rows [(cast x (sym lb" in s and "(cast x (sym ub" in s


def test_group_join_lowers_and_compiles():
    """The nested Op/Reduce form lowers to the group-join plan: a build
    selection, a probe scan with a guarded hash probe, and a per-group fold."""
    text = runtime.compile_check(workloads.group_join_c4())
    assert "group join: build table input 0, probe table input 1" in text
    assert "group field 0 f64 add of payload 0" in text
    assert "group field 1 i64 add of (const i64 1)" in text
    # build scan, probe scan against a rank map and against a hash index, fold,
    # and the rank-map probe that pre-aggregates in LDS (few groups)
    assert text.count("--- kernel ---") == 5
    assert "enum { LDS_AGG = 1, NSLOT = 2 };" in text
    assert "void rf_gj_lds_done(" in text
    assert "? rf_probe(" in text  # probed only where the probe predicate holds
    assert runtime.compile_check(workloads.group_join_c4(), nullable=True)


def test_nested_two_iterator_op_compiles():
    import ref_queries as rq
    text = runtime.compile_check(rq.nested_op())
    assert "scan range rows [(const x 0), (mul" in text


def test_run_reduction_compiles():
    import ref_queries as rq
    text = runtime.compile_check(rq.multidim_op())
    assert "run reduction over the run-end column 0 of table input 0" in text
    assert "rf_seg_fold" in text


def test_union_space_lowers_to_a_merge():
    """An Op over _in(t, A) | _in(t, B) (the reference's UnionSpace merge
    loop) lowers to the union merge kernels, and so does AlgoTrading's union
    with a shifted side, _in(t + (-50), B) (a ShiftedSpace); unions the
    Reffine pass cannot form (an iterator scaled or shifted by a non-constant)
    and an element read at another expression than its side's iterator are
    refused as unsupported at compile time, not lowered wrongly."""
    import ref_queries as rq
    from reffine_amd.reffiner import _add, _func, _get, _i64, _i64_t, _in, _mul, _op, _sub, _sym, _vec_t
    text = runtime.compile_check(rq.micro_join(outer=True))
    assert "union merge of table inputs 0 and 1 (keys ascending)" in text
    assert "rf_qu_count" in text and "rf_qu_emit" in text
    text = runtime.compile_check(workloads.algotrading())
    assert "union merge of table inputs 0 and 1 (keys ascending, shifted by 0 and -50)" in text
    t = _sym("t", _i64_t)
    x = _sym("x", _i64_t)
    A = _sym("left", _vec_t(1, _i64_t, _i64_t))
    B = _sym("right", _vec_t(1, _i64_t, _i64_t))

    def union(side_b, read_b):
        op = _op([t], _in(t, A) | _in(side_b, B), [_sub(_get(A[t], 0), _get(B[read_b], 0))])
        o = _sym("op", op)
        fn = _func("u", o, [A, B, x])
        fn.tbl[o] = op
        return fn
    for side_b, read_b, why in [(_mul(t, _i64(2)), t, "union of iteration spaces"),
                                (_add(t, x), t, "union of iteration spaces"),
                                (_add(t, _i64(3)), t, "other than its _in"),
                                (t, _add(t, _i64(3)), "other than its _in")]:
        with pytest.raises(runtime.ReffineError) as e:
            runtime.compile_check(union(side_b, read_b))
        assert e.value.status == 2, e.value  # RF_ERR_UNSUPPORTED
        assert why in str(e.value)


def test_output_struct_decoder_matches_c_layout():
    """The precompiled output decoder reads each field at its C offset
    (natural alignment, declaration order: ScalarPass,
    src/pass/scalarpass.cpp:103-227) -- checked against a field-by-field
    numpy read on random bytes for the workloads' nested output structs."""
    import numpy as np
    from reffine_amd.runtime import _NP_OF, _Decoder, _layout
    from reffine_amd import workloads as wl

    def slow(buf, t, off=0):
        if t.is_struct():
            out = []
            for d in t.dtypes:
                s, a = _layout(d)
                off = (off + a - 1) // a * a
                out.append(slow(buf, d, off))
                off += s
            return tuple(out)
        if t.is_vector():
            return int.from_bytes(buf[off:off + 8], "little")
        return np.frombuffer(buf, dtype=_NP_OF[t.btype], count=1, offset=off)[0].item()

    rng = np.random.default_rng(4)
    for fn in (wl.group_agg_c3(), wl.scan_select_c2(), wl.scan_agg_c1(), wl.q6_template()):
        t = fn.output.type
        size, _ = _layout(t)
        dec = _Decoder(t)
        dec.vec = [False] * len(dec.vec)  # table handles compared as integers
        for _ in range(10):
            buf = rng.integers(0, 256, size, dtype=np.uint8).tobytes()
            assert repr(dec(buf)) == repr(slow(buf, t))


def test_c_host_program_compiles_against_the_header():
    """tests/c_abi/rf_c_scan.c, a C host of the ABI, compiles against
    include/reffine_hip.h alone (C11, warnings as errors)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c11", "-Wall", "-Werror",
                        "-I", os.path.join(root, "include"),
                        os.path.join(root, "tests", "c_abi", "rf_c_scan.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("name", ["q3", "q4", "q12", "q18", "example"])
def test_tpch_run_stage_queries_compile(name):
    """TPC-H Q3/Q4/Q12/Q18 over the run-end lineitem (benchmark/include/
    tpch.h) lower to a run stage -- a fold per lineitem run, then a scan of
    the runs table with the orders (and customer) probes -- and compile."""
    from reffine_amd import tpch
    text = runtime.compile_check(tpch.QUERIES[name]())
    assert "run stage: the runs table" in text
    assert text.count("--- kernel ---") == 2
    assert "probe 0: table input" in text
    if name == "q12":  # the fold reads orders at the run's key
        assert "probe 0: table input 1 at key (sym $key i64)" in text
    if name == "example":  # supplier probed by each partsupp row inside the fold
        assert "probe 0: table input 1 at key (sym $col.2.1 i64)" in text
        assert "rf_probe(L.probe[0], (int64_t)sm1[row - blo])" in text
    if name == "q3":  # _locate / _readdata into customer
        assert "probe 1: table input 2 key (sym $pcol.0.1 i64)" in text
        assert "(sym $pcol.1.6 i8)" in text


def test_run_table_used_beyond_its_runs_is_rejected():
    """A run-end table read other than through In(t, T) and T[t]
    reductions has no runs-table form."""
    from reffine_amd import tpch
    from reffine_amd.reffiner import _f64, _func, _get, _gt, _in, _len, _op, _red, _sym, _i64_t, _lt, _cast, _idx
    lineitem = _sym("lineitem", tpch.LINEITEM_TYPE)
    orders = _sym("orders", tpch.ORDERS_TYPE)
    k = _sym("k", _i64_t)
    red = _red(lineitem[k], lambda: _f64(0), lambda s, v: s + _get(v, 3))
    rs = _sym("red", red)
    op = _op([k], _in(k, lineitem) & _in(k, orders) & _lt(_idx(0), _len(lineitem, 1)), [rs])
    os_ = _sym("op", op)
    fn = _func("bad", os_, [lineitem, orders])
    fn.tbl[os_] = op
    fn.tbl[rs] = red
    with pytest.raises(runtime.ReffineError):
        runtime.compile_check(fn)


def test_pagerank_materializes_the_indexed_op():
    """The PageRank step (misc.h:212-291): the _buildidx'd Op is split out as
    a function of its own (a run stage over edges) whose result is an
    indexed input of the rest (a run fold over rev_edges probing it)."""
    from reffine_amd import graph
    text = runtime.compile_check(graph.pagerank(1000))
    assert "materialized as an input with an index" in text
    assert text.count("--- kernel ---") == 3
    assert "probe 0: table input 3 at key (sym $col.1.1 i64)" in text


def test_nbody_splits_into_three_plans():
    """Nbody (misc.h:63-210): the pair table and the force table are
    materialized (nested), the force sums are a ranged fold over the pair
    table's windows, the last Op reads both tables at computed rows."""
    from reffine_amd import graph
    text = runtime.compile_check(graph.nbody(64))
    assert text.count("materialized as an input") == 2
    assert "ranged fold: for $key in" in text
    assert "readdata" in text


def test_self_referencing_computed_vector_is_rejected():
    """A computed vector whose definition reads itself is refused (a status
    error), not recursed into without end."""
    from reffine_amd.reffiner import _func, _get, _i64_t, _in, _op, _sym, _vec_t
    A = _sym("A", _vec_t(1, _i64_t, _i64_t))
    t = _sym("t", _i64_t)
    V = _sym("V", _vec_t(1, _i64_t, _i64_t))
    op = _op([t], _in(t, A) & _in(t, V), [_get(A[t], 0)])
    fn = _func("selfref", V, [A])
    fn.tbl[V] = op
    u = _sym("u", _i64_t)
    top = _op([u], _in(u, A) & _in(u, V), [_get(V[u], 0)])
    ts = _sym("top", top)
    fn2 = _func("selfref2", ts, [A])
    fn2.tbl[V] = op
    fn2.tbl[ts] = top
    with pytest.raises(runtime.ReffineError):
        runtime.compile_check(fn2)


def test_independent_components_compile_as_separate_queries():
    """Components that cannot share a scan (two tables; a sequential fold
    beside a sum) lower to one query per component."""
    import seq_queries as sq
    for fn in (sq.sum_and_count_of_two_tables(), sq.decay_and_sum(), sq.selection_and_decay()):
        text = runtime.compile_check(fn)
        assert "independent components, each a query of its own" in text
        assert text.count("\nfunc ") >= 2
