"""The reference test suite's queries (test/src/test_reduce.cpp,
test_transform.cpp, test_join.cpp, test_utils.cpp) through the GPU path,
against the values those tests assert (tests/golden/expected.json)."""
import numpy as np
import pytest

import ref_queries as rq
from oracle import merge
from reffine_amd import DeviceTable, compile_op

pytestmark = pytest.mark.gpu


def students_table(dim=1):
    return DeviceTable.from_columns(rq.students(), dim=dim, names=rq.STUDENT_COLS)


@pytest.mark.parametrize("vectorize", [False, True])
@pytest.mark.parametrize("det", [False, True])
def test_aggregate_golden(device, vectorize, det):
    exp = rq.expected()
    t2 = students_table(2)
    assert compile_op(rq.sum_hours_subvec(), vectorize, det)(t2) == exp["aggregate_loop_test"]
    t1 = students_table(1)
    assert compile_op(rq.vector_op(), vectorize, det)(t1) == exp["aggregate_op_test"]


@pytest.mark.parametrize("vectorize", [False, True])
def test_transform_golden(device, vectorize):
    exp = rq.expected()
    ids, hours = rq.students()[0], rq.students()[1]
    out = compile_op(rq.transform_op(), vectorize)(students_table())
    assert out.nrows == exp["transform_op_test_len"]
    np.testing.assert_array_equal(out.column(0), ids[5:])
    np.testing.assert_array_equal(out.column(1), hours[5:] + 10)


def test_gen_table_and_join_golden(device):
    exp = rq.expected()
    gen = compile_op(rq.gen_table_op())
    tabs = []
    for lb, ub in [(0, 10), (5, 15), (7, 12)]:
        t = gen(lb, ub)
        assert t.nrows == ub - lb
        np.testing.assert_array_equal(t.column(0), np.arange(lb, ub))
        np.testing.assert_array_equal(t.column(1), np.arange(lb, ub))
        tabs.append(t)
    out = compile_op(rq.join_op())(*tabs)
    np.testing.assert_array_equal(out.column(0), exp["join_op_test_keys"])
    np.testing.assert_array_equal(out.column(1), 3 * np.array(exp["join_op_test_keys"]))


def test_empty_range_and_single_row(device):
    gen = compile_op(rq.gen_table_op())
    assert gen(5, 5).nrows == 0
    assert gen(9, 3).nrows == 0
    t = gen(-3, -2)
    np.testing.assert_array_equal(t.column(0), [-3])


@pytest.mark.parametrize("a_ub,b_ub", [(10, 10), (3, 7), (1, 1), (0, 5)])
def test_nested_golden(device, a_ub, b_ub):
    """test/src/test_nested.cpp:42-53: row idx holds (idx / b_ub, idx % b_ub,
    their sum), a outer and b inner."""
    out = compile_op(rq.nested_op(a_ub, b_ub))()
    n = a_ub * b_ub
    assert out.nrows == n
    idx = np.arange(n)
    assert np.array_equal(out.column(0), idx // max(b_ub, 1))
    assert np.array_equal(out.column(1), idx % max(b_ub, 1))
    assert np.array_equal(out.column(2), idx // max(b_ub, 1) + idx % max(b_ub, 1))


def test_multidim_golden(device):
    """test/src/test_multidim.cpp:32-54: one output row per run of the
    run-end encoded x, holding the run value and the sum of d over its rows."""
    e = rq.expected()["multidim_op_test"]
    t = DeviceTable.from_arrow(rq.runend_table(), dim=2)
    out = compile_op(rq.multidim_op())(t)
    assert out.nrows == len(e["keys"])
    assert list(out.column(0)) == e["keys"]
    assert list(out.column(1)) == e["sums"]


def test_run_reduction_random(device):
    """Random runs (lengths 1..200), nulls in the summed column's
    predicate column, int64 and float64 folds: counts and int sums exact,
    the f64 sum bit-exact against the sequential fold of each run."""
    import pyarrow as pa
    import pyarrow.compute as pc
    from reffine_amd.reffiner import (_add, _cast, _f64, _f64_t, _func, _get, _i64, _i64_t, _in,
                                      _new, _op, _red, _sym, _vec_t)
    rng = np.random.default_rng(12)
    lens = rng.integers(1, 200, size=800)
    keys = np.cumsum(rng.integers(1, 5, size=len(lens))).astype(np.int64)
    x = np.repeat(keys, lens)
    n = len(x)
    y = rng.integers(-1000, 1000, size=n).astype(np.int64)
    v = rng.random(n)
    valid_y = rng.random(n) > 0.1
    batch = pa.RecordBatch.from_arrays(
        [pc.run_end_encode(pa.array(x)), pa.array(y, mask=~valid_y), pa.array(v)],
        names=["x", "y", "v"])
    T = _sym("T", _vec_t(2, _i64_t, _i64_t, _f64_t, encodings=["runend", "flat", "flat"]))
    t = _sym("t", _i64_t)
    red = _red(T[t], lambda: _new([_i64(0), _i64(0), _f64(0)]),
               lambda s, r: _new([_add(_get(s, 0), _i64(1)), _add(_get(s, 1), _get(r, 0)),
                                  _add(_get(s, 2), _get(r, 1))]))
    rs = _sym("agg", red)
    op = _op([t], _in(t, T), [_get(rs, 0), _get(rs, 1), _get(rs, 2)])
    os_ = _sym("op", op)
    fn = _func("runs", os_, [T])
    fn.tbl[os_] = op
    fn.tbl[rs] = red
    out = compile_op(fn)(DeviceTable.from_arrow(batch, dim=2))
    ends = np.cumsum(lens)
    starts = ends - lens
    assert np.array_equal(out.column(0), keys)
    cnt, sy, sv = out.column(1), out.column(2), out.column(3)
    for r in range(len(lens)):
        m = valid_y[starts[r]:ends[r]]  # SubVector rows qualify by column 1's validity
        assert cnt[r] == m.sum()
        assert sy[r] == y[starts[r]:ends[r]][m].sum()
        acc = 0.0
        for val in v[starts[r]:ends[r]][m]:
            acc = acc + float(val)
        assert np.float64(sv[r]).view(np.int64) == np.float64(acc).view(np.int64)


def test_micro_benchmark_programs(device):
    """benchmark/include/micro.h: select (even keys), inner join (keys of
    both tables, value difference), sum, and the outer join (UnionSpace)
    against the oracle's merge."""
    rng = np.random.default_rng(3)
    lk = np.unique(rng.integers(0, 400_000, 150_000)).astype(np.int64)
    rk = np.unique(rng.integers(0, 400_000, 150_000)).astype(np.int64)
    lv = rng.integers(-10**9, 10**9, len(lk)).astype(np.int64)
    rv = rng.integers(-10**9, 10**9, len(rk)).astype(np.int64)
    L = DeviceTable.from_columns([lk, lv], dim=1).build_index()
    R = DeviceTable.from_columns([rk, rv], dim=1).build_index()
    sel = compile_op(rq.micro_select())(L)
    np.testing.assert_array_equal(sel.column(0), lk[lk % 2 == 0])
    inner = compile_op(rq.micro_join())(L, R)
    both, li, ri = np.intersect1d(lk, rk, return_indices=True)
    np.testing.assert_array_equal(inner.column(0), both)
    np.testing.assert_array_equal(inner.column(1), lv[li] - rv[ri])
    L2 = DeviceTable.from_columns([lk, lv], dim=2)
    assert compile_op(rq.micro_sum())(L2) == int(lv.sum())
    outer = compile_op(rq.micro_join(outer=True))(L, R)
    t, d = merge.ojoin(lk, lv, rk, rv)
    np.testing.assert_array_equal(outer.column(0), t)
    np.testing.assert_array_equal(outer.column(1), d)


@pytest.mark.gpu
@pytest.mark.parametrize("det", [False, True])
def test_float32_sum_exact(device, det):
    """A float32 SubVector sum (micro.h:104-115's shape): the wave reduction
    moves float partials between lanes as bits. Quarters sum exactly in any
    order below 2^22, so the result is exact; a value conversion of the
    partials at any butterfly level would drop their fractions."""
    from reffine_amd.reffiner import (_add, _f32, _f32_t, _func, _get, _i64_t, _idx, _len,
                                      _red, _subvec, _sym, _vec_t)
    vec_in = _sym("vec_in", _vec_t(2, _i64_t, _f32_t))
    red = _red(_subvec(vec_in, _idx(0), _len(vec_in, 1)), lambda: _f32(0.0),
               lambda s, v: _add(s, _get(v, 0)))
    rs = _sym("red", red)
    fn = _func("sum_f32", rs, [vec_in])
    fn.tbl[rs] = red
    q = compile_op(fn, deterministic=det)
    for n in (1, 63, 4096, 100_003):
        x = np.full(n, 0.25, dtype=np.float32)
        got = q(DeviceTable.from_columns([np.arange(n, dtype=np.int64), x], dim=2))
        assert float(got) == 0.25 * n, (n, got)
    y = np.random.default_rng(11).integers(0, 64, 50_000).astype(np.float32) * np.float32(0.75)
    yk = np.arange(len(y), dtype=np.int64)
    assert float(q(DeviceTable.from_columns([yk, y], dim=2))) == float(y.astype(np.float64).sum())
