"""TPC-H Q3, Q4, Q12 and Q18 as reffine writes them (benchmark/include/
tpch.h) over a run-end encoded lineitem, through the GPU run stage, against
oracle.tpch on the same seeded tables: keys, counts and FP64 sums
bit-exact (each run is folded in row order, as the reference's loop does).

The tables exercise every join path: orders missing from the orders table
(their lines skip the merge), orders with no lines, and orders whose
customer key is absent (Q3's _locate returns -1)."""
import numpy as np
import pytest

from oracle import tpch as ot
from reffine_amd import DeviceTable, compile_op, datagen, tpch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    t = datagen.tpch_tables(20000, seed=5)
    rb = tpch.record_batches(t)
    dev = {"lineitem": DeviceTable.from_arrow(rb["lineitem"], dim=2),
           "orders": DeviceTable.from_arrow(rb["orders"]).build_index(),
           "customer": DeviceTable.from_arrow(rb["customer"]).build_index()}
    return t, dev


def _args(dev, name):
    return [dev[a] for a in tpch.ARGS[name]]


@pytest.mark.parametrize("quantity", [300, 200, 120])
def test_q18(device, tables, quantity):
    t, dev = tables
    out = compile_op(tpch.q18(quantity))(*_args(dev, "q18"))
    k, v = ot.q18(t, quantity)
    if quantity == 120:
        assert len(k) > 1000
    assert out.nrows == len(k)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1).view(np.uint64), v.view(np.uint64))


@pytest.mark.parametrize("segment,date", [(1, 795484800), (3, 795484800), (0, 760000000)])
def test_q3(device, tables, segment, date):
    t, dev = tables
    out = compile_op(tpch.q3(segment, date))(*_args(dev, "q3"))
    k, v = ot.q3(t, segment, date)
    assert len(k) > 20
    assert out.nrows == len(k)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1).view(np.uint64), v.view(np.uint64))


@pytest.mark.parametrize("start,end", [(700000000, 900000000), (760000000, 780000000)])
def test_q4(device, tables, start, end):
    t, dev = tables
    got = compile_op(tpch.q4(start, end))(*_args(dev, "q4"))
    assert list(got) == ot.q4(t, start, end)


def test_q12(device, tables):
    t, dev = tables
    out = compile_op(tpch.q12())(*_args(dev, "q12"))
    k, h, lo = ot.q12(t)
    assert out.nrows == len(k)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1), h)
    np.testing.assert_array_equal(out.column(2), lo)
    assert h.sum() > 0 and lo.sum() > 0


def test_run_stage_repeats_and_times(device, tables):
    """A second call reuses the compiled kernels and gives the same table;
    Query.time covers both stages."""
    t, dev = tables
    q = compile_op(tpch.q18(150))
    a = q(*_args(dev, "q18"))
    b = q(*_args(dev, "q18"))
    np.testing.assert_array_equal(a.column(0), b.column(0))
    np.testing.assert_array_equal(a.column(1), b.column(1))
    assert q.time(*_args(dev, "q18"), iters=3) > 0


def test_small_tables(device):
    """Tiny tables (a handful of orders, one line each or none) through the
    same plans."""
    t = datagen.tpch_tables(9, seed=3)
    rb = tpch.record_batches(t)
    dev = {"lineitem": DeviceTable.from_arrow(rb["lineitem"], dim=2),
           "orders": DeviceTable.from_arrow(rb["orders"]).build_index(),
           "customer": DeviceTable.from_arrow(rb["customer"]).build_index()}
    out = compile_op(tpch.q18(0))(*_args(dev, "q18"))
    k, v = ot.q18(t, 0)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1), v)
    assert list(compile_op(tpch.q4())(*_args(dev, "q4"))) == ot.q4(t)


def _locate_fn(T_type, P_type):
    from reffine_amd.reffiner import _func, _get, _in, _locate, _op, _sym, _i64_t
    T = _sym("T", T_type)
    P = _sym("P", P_type)
    t = _sym("t", _i64_t)
    op = _op([t], _in(t, P), [_locate(T, _get(P[t], 0))])
    os_ = _sym("op", op)
    fn = _func("locate", os_, [T, P])
    fn.tbl[os_] = op
    return fn


@pytest.mark.parametrize("form", ["ascending", "shuffled_dups", "sparse", "nulls"])
def test_locate_index_forms(device, form):
    """_locate (vector_locate, include/reffine/vinstr/vinstr.h:137-146) over
    each form of the table index: the rank map (keys dense in their range,
    strictly ascending or not, with repeated keys: the first row) and the
    hash slots (sparse keys, or null keys, which are never found)."""
    import pyarrow as pa
    from reffine_amd.reffiner import _i64_t, _vec_t
    rng = np.random.default_rng(7)
    n = 50_000
    valid = None
    if form == "ascending":
        keys = np.arange(n, dtype=np.int64) * 3 + 1000
    elif form == "shuffled_dups":
        keys = rng.integers(-500, n // 2, n).astype(np.int64)
    elif form == "sparse":
        keys = rng.integers(-2**62, 2**62, n).astype(np.int64)
        keys[:100] = keys[100:200]
    else:
        keys = rng.integers(0, n, n).astype(np.int64)
        valid = rng.random(n) > 0.2
    first = {}
    for r, k in enumerate(keys.tolist()):
        if (valid is None or valid[r]) and k not in first:
            first[k] = r
    probe = np.concatenate([keys[rng.integers(0, n, 4000)], rng.integers(-10**6, 10**6, 4000),
                            keys[:50]]).astype(np.int64)
    T = DeviceTable.from_arrow(pa.RecordBatch.from_arrays(
        [pa.array(keys, mask=None if valid is None else ~valid), pa.array(np.zeros(n))],
        names=["k", "x"]))
    T.build_index()
    P = DeviceTable.from_arrow(pa.RecordBatch.from_arrays(
        [pa.array(np.arange(len(probe), dtype=np.int64)), pa.array(probe)], names=["i", "k"]))
    from reffine_amd.reffiner import _f64_t
    out = compile_op(_locate_fn(_vec_t(1, _i64_t, _f64_t), _vec_t(1, _i64_t, _i64_t)))(T, P)
    want = np.array([first.get(k, -1) for k in probe.tolist()], dtype=np.int64)
    np.testing.assert_array_equal(out.column(1), want)
