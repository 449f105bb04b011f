"""Sequential folds on the GPU (folds the scan cannot re-associate, run by
one wave in row order -- the reference's loop, src/pass/loopgen.cpp:
253-262) against row loops and the CPU JIT restatement, bit for bit; and
per-row reads of a run-end encoded column (the run value at the row index,
src/pass/readwritepass.cpp:16-29) in both the scan and the sequential fold."""
import numpy as np
import pytest

import seq_queries as sq
from oracle import cjit
from reffine_amd import DeviceTable, compile_op
from reffine_amd.reffiner import (_add, _f64, _f64_t, _func, _get, _i64_t, _idx, _len, _mul,
                                  _red, _subvec, _sym, _vec_t)

pytestmark = pytest.mark.gpu


def _data(name, n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(-1000, 1000, n).astype(np.int64)
    v = rng.random(n)
    if name == "maxv":
        v[rng.random(n) < 0.05] = np.nan
    return k, v


def _bits(t):
    return [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in t]


@pytest.mark.parametrize("name", sorted(sq.BUILDERS))
@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 100_003])
def test_sequential_fold_matches_row_loop(device, name, n):
    k, v = _data(name, n, seed=n + 3)
    t = DeviceTable.from_columns([np.arange(n, dtype=np.int64), k, v], dim=2)
    q = compile_op(sq.BUILDERS[name]())
    assert "sequential fold" in q.describe()
    got = q(t)
    got = got if isinstance(got, tuple) else (got,)
    want = sq.loop(name, k, v)
    assert _bits(got) == _bits(want)
    if n:
        (cj,) = cjit.compile_op(sq.BUILDERS[name]())({"cols": [np.arange(n), k, v]})
        assert _bits(got) == _bits(cj)
    # the same bits in the deterministic mode and on a second run
    qd = compile_op(sq.BUILDERS[name](), deterministic=True)
    gd = qd(t)
    assert _bits(gd if isinstance(gd, tuple) else (gd,)) == _bits(want)


@pytest.mark.parametrize("name", ["decay", "pair"])
def test_sequential_fold_skips_null_rows(device, name):
    n = 20_000
    k, v = _data(name, n, seed=9)
    valid = np.random.default_rng(2).random(n) > 0.2
    t = DeviceTable.from_columns([np.arange(n, dtype=np.int64), k, v], dim=2,
                                 validity=[None, valid, None])
    got = compile_op(sq.BUILDERS[name]())(t)
    got = got if isinstance(got, tuple) else (got,)
    assert _bits(got) == _bits(sq.loop(name, k, v, valid))


def test_sequential_fold_in_a_batch_and_served(device):
    """A batch or serving unit of a sequential fold runs on the plain path."""
    from reffine_amd.batch import Batch, serve
    n = 50_000
    k, v = _data("decay", n, seed=4)
    t = DeviceTable.from_columns([np.arange(n, dtype=np.int64), k, v], dim=2)
    q = compile_op(sq.decay())
    res = Batch().run([(q, (t,)), (q, (t,))])
    assert _bits(tuple(res[0]) if isinstance(res[0], tuple) else (res[0],)) == \
        _bits(sq.loop("decay", k, v))
    assert (serve([(q, (t,), 0.0)] * 4).status == 0).all()


REE_T = _vec_t(2, _i64_t, _f64_t, encodings=["flat", "runend"])


def _ree_table(n, seed):
    import pyarrow as pa
    import pyarrow.compute as pc
    rng = np.random.default_rng(seed)
    y = np.repeat(rng.random(n // 4 + 1), 4)[:n]  # runs of 4 rows
    rb = pa.RecordBatch.from_arrays(
        [pa.array(np.arange(n, dtype=np.int64)),
         pc.run_end_encode(pa.array(y), run_end_type=pa.int32())], names=["rowid", "y"])
    runs = y[::4]
    return DeviceTable.from_arrow(rb, dim=2), runs


def _ree_fold(kind):
    T = _sym("T", REE_T)
    if kind == "sum":
        acc = lambda s, r: _add(s, _get(r, 0))  # noqa: E731
    else:
        acc = lambda s, r: _add(_mul(s, _f64(0.5)), _get(r, 0))  # noqa: E731
    red = _red(_subvec(T, _idx(0), _len(T, 0)), lambda: _f64(0), acc)
    rs = _sym("r", red)
    fn = _func("ree_" + kind, rs, [T])
    fn.tbl[rs] = red
    return fn


@pytest.mark.parametrize("kind", ["sum", "decay"])
def test_run_end_column_read_per_row(device, kind):
    """A fold over rows [0, rows) of a table whose value column is run-end
    encoded reads, at row i, the run value of index i (the values child at
    the row index, as the reference's ReadData of a run-end child does); rows
    past the number of runs read 0 here."""
    n = 10_000
    t, runs = _ree_table(n, 5)
    got = compile_op(_ree_fold(kind))(t)
    vals = np.zeros(n)
    vals[:len(runs)] = runs
    if kind == "sum":
        assert abs(got - float(np.sum(vals))) <= 1e-12 * abs(float(np.sum(vals)))
    else:
        s = 0.0
        for x in vals:
            s = s * 0.5 + float(x)
        assert np.float64(got).view(np.uint64) == np.float64(s).view(np.uint64)
