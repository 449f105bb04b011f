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


def test_run_stage_nulls_and_a_second_table(device):
    """A run stage over random runs with nulls in column 1 (the SubVector
    body condition: null rows are not folded), merged with a second table
    that lacks some keys; the runs' sums filtered and output with a column
    of the second table. Expected values from the sequential fold of each
    run."""
    import pyarrow as pa
    import pyarrow.compute as pc
    from reffine_amd.reffiner import (_add, _f64, _f64_t, _func, _get, _gt, _i64, _i64_t, _in,
                                      _new, _op, _red, _sym, _vec_t)
    rng = np.random.default_rng(31)
    lens = rng.integers(1, 40, size=3000)
    keys = np.cumsum(rng.integers(1, 4, size=len(lens))).astype(np.int64)
    x = np.repeat(keys, lens)
    n = len(x)
    y = rng.integers(-50, 50, size=n).astype(np.int64)
    v = rng.random(n)
    valid_y = rng.random(n) > 0.1
    T = DeviceTable.from_arrow(pa.RecordBatch.from_arrays(
        [pc.run_end_encode(pa.array(x)), pa.array(y, mask=~valid_y), pa.array(v)],
        names=["x", "y", "v"]), dim=2)
    ukeys = keys[rng.random(len(keys)) > 0.3]
    w = rng.integers(0, 1000, size=len(ukeys)).astype(np.int64)
    U = DeviceTable.from_columns([ukeys, w], dim=1).build_index()
    Ts = _sym("T", _vec_t(2, _i64_t, _i64_t, _f64_t, encodings=["runend", "flat", "flat"]))
    Us = _sym("U", _vec_t(1, _i64_t, _i64_t))
    t = _sym("t", _i64_t)
    red = _red(Ts[t], lambda: _new([_i64(0), _f64(0)]),
               lambda s, r: _new([_add(_get(s, 0), _get(r, 0)), _add(_get(s, 1), _get(r, 1))]))
    rs = _sym("agg", red)
    op = _op([t], _in(t, Ts) & _in(t, Us) & _gt(_get(rs, 0), _i64(0)),
             [_get(rs, 0), _get(rs, 1), _get(Us[t], 0)])
    os_ = _sym("op", op)
    fn = _func("runs2", os_, [Ts, Us])
    fn.tbl[os_] = op
    fn.tbl[rs] = red
    out = compile_op(fn)(T, U)
    wmap = dict(zip(ukeys.tolist(), w.tolist()))
    ends = np.cumsum(lens)
    starts = ends - lens
    ek, es, ev, ew = [], [], [], []
    for r in range(len(lens)):
        k = int(keys[r])
        if k not in wmap:
            continue
        sy, sv = 0, 0.0
        for i in range(starts[r], ends[r]):
            if not valid_y[i]:
                continue  # the SubVector loop skips the row
            sy += int(y[i])
            sv = sv + float(v[i])
        if sy > 0:
            ek.append(k); es.append(sy); ev.append(sv); ew.append(wmap[k])
    assert len(ek) > 200
    np.testing.assert_array_equal(out.column(0), ek)
    np.testing.assert_array_equal(out.column(1), es)
    np.testing.assert_array_equal(out.column(2).view(np.uint64), np.array(ev).view(np.uint64))
    np.testing.assert_array_equal(out.column(3), ew)


def test_run_stage_deterministic_and_batched(device, tables):
    """Deterministic compilation gives the same results; in a batch (and
    served), run-stage units run on the plain path beside the round's scan
    units and return the same outputs."""
    from reffine_amd.batch import Batch, serve
    from reffine_amd.reffiner import _f64, _func, _red, _subvec, _sym, _idx, _len, _add, _get
    t, dev = tables
    assert list(compile_op(tpch.q4(), deterministic=True)(*_args(dev, "q4"))) == ot.q4(t)
    out = compile_op(tpch.q18(150), deterministic=True)(*_args(dev, "q18"))
    k, v = ot.q18(t, 150)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1), v)
    q18 = compile_op(tpch.q18(150), deterministic=True)
    q4 = compile_op(tpch.q4(), deterministic=True)
    from reffine_amd.reffiner import _vec_t, _i64_t, _i8_t, _f64_t, _i32_t
    o = _sym("orders", _vec_t(2, _i64_t, _i64_t, _i8_t, _f64_t, _i64_t, _i8_t, _i64_t, _i32_t,
                              _i64_t))
    tot = _red(_subvec(o, _idx(0), _len(o, 1)), lambda: _f64(0), lambda s, r: _add(s, _get(r, 2)))
    ts = _sym("tot", tot)
    fn = _func("total_price", ts, [o])
    fn.tbl[ts] = tot
    qs = compile_op(fn, deterministic=True)
    o2 = dev["orders"].view(2)
    work = [(q18, tuple(_args(dev, "q18"))), (qs, (o2,)), (q4, tuple(_args(dev, "q4")))]
    res = Batch().run(work)  # a table output comes back as its host columns
    np.testing.assert_array_equal(res[0][0], k)
    np.testing.assert_array_equal(res[0][1], v)
    assert res[1] == qs(o2)
    assert list(res[2]) == ot.q4(t)
    sr = serve([(q4, tuple(_args(dev, "q4")), 0.0), (qs, (o2,), 0.0)] * 3)
    assert (sr.status == 0).all()


def test_plain_batch_unit_table_output_needs_out_table(device, tables):
    """A run-stage unit (plain path inside the batch) whose output is a
    table: rf_batch_result with only output_addr is refused (the batch's own
    table is never handed out), and with out_table it gives a view that
    stays valid after the reset, freed by the caller."""
    import ctypes
    from reffine_amd.batch import Batch
    from reffine_amd.runtime import lib
    _, dev = tables
    q18 = compile_op(tpch.q18(150))
    want = q18(*_args(dev, "q18"))
    b = Batch()
    u = b.add(q18, *_args(dev, "q18"))
    b.launch()
    b.wait()
    buf = ctypes.create_string_buffer(16)
    assert lib().rf_batch_result(b.handle, u, buf, None) == 1  # RF_ERR_INVALID
    assert b"out_table" in lib().rf_last_error()
    tbl = ctypes.c_void_p()
    assert lib().rf_batch_result(b.handle, u, buf, ctypes.byref(tbl)) == 0
    assert ctypes.c_void_p.from_buffer(buf).value == tbl.value
    b.reset()
    view = DeviceTable(tbl.value)
    np.testing.assert_array_equal(view.column(0), want.column(0))
    np.testing.assert_array_equal(view.column(1), want.column(1))
    view.free()


def test_run_stage_empty_lineitem(device, tables):
    """No lineitem rows: no runs, empty outputs and zero counts."""
    t0 = datagen.tpch_tables(5, seed=2)
    t0["lineitem"] = {c: a[:0] for c, a in t0["lineitem"].items()}
    rb = tpch.record_batches(t0)
    li = DeviceTable.from_arrow(rb["lineitem"], dim=2)
    _, dev = tables
    assert compile_op(tpch.q18(0))(li, dev["orders"]).nrows == 0
    assert list(compile_op(tpch.q4())(li, dev["orders"])) == [0, 0, 0, 0, 0]


@pytest.mark.parametrize("nation,size", [(0, 15), (3, 40), (24, 0)])
def test_example(device, nation, size):
    """TPCHQueryExample (tpch.h:606-682): the fold over a part's partsupp
    run probes supplier by each row's supplier key."""
    t = datagen.tpch_parts(20000, seed=4)
    rb = tpch.part_batches(t)
    dev = {"part": DeviceTable.from_arrow(rb["part"]).build_index(),
           "supplier": DeviceTable.from_arrow(rb["supplier"]).build_index(),
           "partsupp": DeviceTable.from_arrow(rb["partsupp"], dim=2)}
    out = compile_op(tpch.example(nation, size))(*[dev[a] for a in tpch.ARGS["example"]])
    k, v = ot.example(t, nation, size)
    assert len(k) > 1000 and v.sum() > 0
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1), v)


def _close(got, want, rel=1e-12):
    return abs(got - want) <= rel * max(abs(want), 1e-300)


def test_q1_and_q6(device, tables):
    """Q1 and Q6 (tpch.h:160-216, 398-486) as written, over the same
    lineitem: Q1 folds the first (number of runs) rows, the reference's
    _len of its run-end column 0."""
    t, dev = tables
    got = compile_op(tpch.q1())(dev["lineitem"])
    want = ot.q1(t)
    flat = [x for x in (got if isinstance(got, (list, tuple)) else [got])]
    assert len(flat) == 15
    for g, w in zip(flat, want):
        assert _close(g, w), (g, w)
    assert sum(want[4::5]) == len(np.unique(t["lineitem"]["l_orderkey"]))
    got6 = compile_op(tpch.q6())(dev["lineitem"])
    assert _close(got6, ot.q6(t))
    assert ot.q6(t) > 0


def test_tpcds_q9(device):
    ss = datagen.store_sales(300_000, seed=7)
    T = DeviceTable.from_arrow(tpch.store_sales_batch(ss), dim=2)
    got = compile_op(tpch.tpcds_q9())(T)
    want = ot.tpcds_q9(ss)
    flat = []
    for x in got:
        flat += list(x) if isinstance(x, (list, tuple)) else [x]
    got = [flat[3 * b:3 * b + 3] for b in range(5)]
    for (c, a, b), (wc, wa, wb) in zip(got, want):
        assert c == wc and _close(a, wa) and _close(b, wb)


def test_q11(device):
    """Q11 as written: its symbol table holds _initval nodes and a second
    Op that the output does not reach."""
    t = datagen.tpch_parts(20000, seed=6)
    rb = tpch.part_batches(t)
    sup = DeviceTable.from_arrow(rb["supplier"]).build_index()
    ps = DeviceTable.from_arrow(rb["partsupp"], dim=2)
    for nation in (0, 5):
        out = compile_op(tpch.q11(nation))(sup, ps)
        k, v = ot.q11(t, nation)
        assert len(k) > 100
        np.testing.assert_array_equal(out.column(0), k)
        np.testing.assert_array_equal(out.column(1).view(np.uint64), v.view(np.uint64))


def test_example_probe_misses(device):
    """The run fold's row-keyed probe over a packed rank map (supplier keys
    strictly ascending, with gaps): partsupp rows name suppliers below the
    key range, past it and in interior gaps. A missing supplier's columns
    read as 0 (its nation is then 0), folded in row order per part."""
    t = datagen.tpch_parts(20000, seed=8)
    sup, ps = t["supplier"], t["partsupp"]
    n_s = len(sup["s_suppkey"])
    i = np.arange(n_s, dtype=np.int64)
    new_keys = 100 + 3 * i + (i % 2)  # strictly ascending, every gap >= 2
    supp = new_keys[ps["ps_suppkey"] - 1]
    r = np.arange(len(supp))
    supp = np.where(r % 17 == 3, 7, supp)                      # below the range
    supp = np.where(r % 19 == 5, new_keys[-1] + 11, supp)       # past the range
    supp = np.where(r % 13 == 2, new_keys[(r * 7) % n_s] + 1, supp)  # interior gaps
    assert not np.isin(supp[r % 13 == 2], new_keys).any()
    t["supplier"] = dict(sup, s_suppkey=new_keys)
    t["partsupp"] = dict(ps, ps_suppkey=supp)
    rb = tpch.part_batches(t)
    dev = {"part": DeviceTable.from_arrow(rb["part"]).build_index(),
           "supplier": DeviceTable.from_arrow(rb["supplier"]).build_index(),
           "partsupp": DeviceTable.from_arrow(rb["partsupp"], dim=2)}
    srow = {int(k): j for j, k in enumerate(new_keys.tolist())}
    nation = t["supplier"]["s_nationkey"]
    for nation_key, size in ((0, 10), (4, 0)):
        out = compile_op(tpch.example(nation_key, size))(
            *[dev[a] for a in tpch.ARGS["example"]])
        keys, vals = [], []
        pk, qty = t["partsupp"]["ps_partkey"], t["partsupp"]["ps_availqty"]
        psize = dict(zip(t["part"]["p_partkey"].tolist(), t["part"]["p_size"].tolist()))
        lo = 0
        while lo < len(pk):
            hi = lo
            while hi < len(pk) and pk[hi] == pk[lo]:
                hi += 1
            v = np.int32(0)
            for j in range(lo, hi):
                row = srow.get(int(supp[j]), -1)
                nk = int(nation[row]) if row >= 0 else 0
                if nk == nation_key:
                    v = np.int32(v + np.int32(qty[j]))
            if psize[int(pk[lo])] > size:
                keys.append(int(pk[lo]))
                vals.append(int(v))
            lo = hi
        assert len(keys) > 1000
        np.testing.assert_array_equal(out.column(0), np.array(keys, np.int64))
        np.testing.assert_array_equal(out.column(1), np.array(vals, np.int32))
