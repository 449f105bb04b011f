"""Arrow ingest of sliced record batches (ArrowArray.offset != 0).

A slice of a batch is exported through the Arrow C data interface with
offsets instead of copies; rf_table_upload must upload exactly the slice's
rows: values, validity bits (also at offsets that are not a whole byte) and,
for a run-end column, the runs that cover the slice, cut to it. The reference
ignores the offset field (include/reffine/vinstr/vinstr.h:25-111), which reads
the wrong rows of a slice; these tests pin the Arrow semantics instead.
"""
from __future__ import annotations

import numpy as np
import pytest

from reffine_amd import DeviceTable, compile_op

pytestmark = pytest.mark.gpu

SLICES = [(0, 10_000), (1, 500), (7, 1000), (8, 3001), (13, 9_987), (10_000, 0), (4095, 17),
          (9_999, 1)]


def _batch(n=10_000, seed=3):
    import pyarrow as pa
    rng = np.random.default_rng(seed)
    cols = {"a": rng.integers(-2**40, 2**40, n).astype(np.int64),
            "b": rng.random(n),
            "c": rng.integers(-1000, 1000, n).astype(np.int32),
            "d": rng.integers(-100, 100, n).astype(np.int8)}
    valid = {"b": rng.random(n) > 0.2, "d": rng.random(n) > 0.5}
    arrs = [pa.array(cols[k], mask=~valid[k]) if k in valid else pa.array(cols[k]) for k in cols]
    return pa.RecordBatch.from_arrays(arrs, names=list(cols)), cols, valid


@pytest.mark.parametrize("form", ["batch", "struct"])
@pytest.mark.parametrize("off,ln", SLICES)
def test_sliced_batch_uploads_its_rows(device, off, ln, form):
    """pyarrow exports a sliced RecordBatch with offsets on the columns and a
    sliced StructArray with the offset on the struct: both upload the slice."""
    import pyarrow as pa
    batch, cols, valid = _batch()
    if form == "batch":
        t = DeviceTable.from_arrow(batch.slice(off, ln))
    else:
        sa = pa.StructArray.from_arrays(batch.columns, names=batch.schema.names)
        t = DeviceTable.from_arrow(sa.slice(off, ln))
    assert t.nrows == ln
    for c, k in enumerate(cols):
        want = cols[k][off:off + ln]
        got = t.column(c)
        assert len(got) == ln
        if k in valid:
            m = valid[k][off:off + ln]
            assert np.array_equal(t.validity(c), m)
            assert np.array_equal(got[m], want[m])
        else:
            assert np.array_equal(got, want)


def test_query_over_a_slice_equals_the_same_rows_uploaded(device):
    """The C1-shaped scan (nullable columns) over a slice at an odd offset
    against the same rows uploaded from fresh arrays."""
    import pyarrow as pa
    from reffine_amd import workloads
    rng = np.random.default_rng(9)
    n = 50_000
    rowid = np.arange(n, dtype=np.int64)
    k = rng.integers(0, 1000, n).astype(np.int64)
    v = rng.random(n)
    d = rng.integers(0, 2556, n).astype(np.int32)
    vv = rng.random(n) > 0.1
    batch = pa.RecordBatch.from_arrays([pa.array(rowid), pa.array(k), pa.array(v, mask=~vv),
                                        pa.array(d)], names=["id", "k", "v", "d"])
    off, ln = 12_345, 30_001
    q = compile_op(workloads.scan_agg_c1())
    got = q(DeviceTable.from_arrow(batch.slice(off, ln), dim=2))
    s = slice(off, off + ln)
    fresh = DeviceTable.from_columns([rowid[s], k[s], v[s], d[s]], dim=2,
                                     validity=[None, None, vv[s], None])
    want = q(fresh)
    assert got[0] == want[0]
    assert np.float64(got[1]).view(np.int64) == np.float64(want[1]).view(np.int64)


@pytest.mark.parametrize("off,ln", [(0, 20_000), (3, 1000), (150, 7_777), (19_999, 1), (777, 0)])
def test_sliced_run_end_column(device, off, ln):
    """A run-end column sliced inside runs: the run reduction over the slice
    sees the slice's runs (the first and last cut), with y's validity bits
    taken at the slice's bit offset."""
    import pyarrow as pa
    import pyarrow.compute as pc
    from reffine_amd.reffiner import (_add, _f64, _f64_t, _func, _get, _i64, _i64_t, _in, _new,
                                      _op, _red, _sym, _vec_t)
    rng = np.random.default_rng(21)
    lens = rng.integers(1, 60, size=700)
    keys = np.cumsum(rng.integers(1, 5, size=len(lens))).astype(np.int64)
    x = np.repeat(keys, lens)[:20_000]
    n = len(x)
    y = rng.integers(-1000, 1000, size=n).astype(np.int64)
    v = rng.random(n)
    vy = rng.random(n) > 0.1
    batch = pa.RecordBatch.from_arrays(
        [pc.run_end_encode(pa.array(x)), pa.array(y, mask=~vy), pa.array(v)],
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
    out = compile_op(fn)(DeviceTable.from_arrow(batch.slice(off, ln), dim=2))
    xs, ys, vs, ms = x[off:off + ln], y[off:off + ln], v[off:off + ln], vy[off:off + ln]
    if ln == 0:
        assert out.nrows == 0
        return
    cut = np.flatnonzero(np.diff(xs)) + 1
    starts = np.concatenate([[0], cut])
    ends = np.concatenate([cut, [ln]])
    assert np.array_equal(out.column(0), xs[starts])
    cnt, sy, sv = out.column(1), out.column(2), out.column(3)
    for r in range(len(starts)):
        m = ms[starts[r]:ends[r]]
        assert cnt[r] == m.sum()
        assert sy[r] == ys[starts[r]:ends[r]][m].sum()
        acc = 0.0
        for val in vs[starts[r]:ends[r]][m]:
            acc = acc + float(val)
        assert np.float64(sv[r]).view(np.int64) == np.float64(acc).view(np.int64)


def _table_with_nulls(n=5000, seed=11):
    rng = np.random.default_rng(seed)
    cols = [np.arange(n, dtype=np.int64), rng.integers(-9, 9, n).astype(np.int64), rng.random(n),
            rng.integers(0, 100, n).astype(np.int32)]
    valid = [None, None, rng.random(n) > 0.3, None]
    return cols, valid


def test_device_export_import_round_trip(device):
    """A table exported as an Arrow C Device array (ROCm, zero copy) and
    imported again borrows the same buffers, keeps them alive after the
    original is freed, and gives the same columns, validity and query
    results."""
    from reffine_amd import runtime, workloads
    cols, valid = _table_with_nulls()
    t = DeviceTable.from_columns(cols, dim=2, validity=valid)
    e = t.export_device()
    assert e.array.device_type == runtime.ARROW_DEVICE_ROCM and e.array.device_id == 0
    assert e.schema.format == b"+s" and e.schema.n_children == 4
    assert e.array.array.length == len(cols[0]) and e.array.sync_event is None
    assert e.schema.children[1].contents.format == b"l"
    assert e.schema.children[2].contents.format == b"g"
    ptr = t.device_ptr(2)
    q = compile_op(workloads.scan_agg_c1(0, 60, 0.9))
    want = q(t)
    u = e.import_as_table(dim=2)
    assert u.device_ptr(2) == ptr  # borrowed, not copied
    t.free()
    del t
    for c in range(4):
        assert np.array_equal(u.column(c), cols[c]) if valid[c] is None else \
            np.array_equal(u.column(c)[valid[c]], cols[c][valid[c]])
    assert np.array_equal(u.validity(2), valid[2])
    got = q(u)
    assert got[0] == want[0]
    assert np.float64(got[1]).view(np.int64) == np.float64(want[1]).view(np.int64)


@pytest.mark.parametrize("off", [16, 3])
def test_device_import_of_a_slice(device, off):
    """A device array whose struct offset keeps the kernels' load granules
    (16 rows of 8 bytes, 16 validity bits) is borrowed; any other offset is
    imported by a copy. Both hold exactly the slice's rows."""
    cols, valid = _table_with_nulls(seed=12)
    n = len(cols[0])
    t = DeviceTable.from_columns(cols, dim=1, validity=valid)
    e = t.export_device()
    e.array.array.offset = off
    e.array.array.length = n - off - 5
    u = e.import_as_table()
    ln = n - off - 5
    assert u.nrows == ln
    s = slice(off, off + ln)
    if off % 16 == 0:
        assert u.device_ptr(1) == t.device_ptr(1) + off * 8
    else:
        assert u.device_ptr(1) != t.device_ptr(1) + off * 8
    assert np.array_equal(u.column(1), cols[1][s])
    assert np.array_equal(u.column(3), cols[3][s])
    assert np.array_equal(u.validity(2), valid[2][s])
    m = valid[2][s]
    assert np.array_equal(u.column(2)[m], cols[2][s][m])


def test_host_batch_through_the_device_interface(device):
    """pyarrow exports a host batch as an Arrow C Device array of device
    type CPU; it is uploaded like rf_table_upload, slices included."""
    batch, cols, valid = _batch(n=3000, seed=5)
    for off, ln in ((0, 3000), (37, 1000)):
        t = DeviceTable.from_arrow_c_device(batch.slice(off, ln))
        assert t.nrows == ln
        assert np.array_equal(t.column(0), cols["a"][off:off + ln])
        assert np.array_equal(t.validity(1), valid["b"][off:off + ln])


def test_run_end_table_exports_and_imports(device):
    """A run-end column exports as a "+r" child (run ends + values) and
    imports again (by a copy: run ends are not borrowed) with the same
    run reduction result."""
    import ref_queries as rq
    t = DeviceTable.from_arrow(rq.runend_table(), dim=2)
    e = t.export_device()
    assert e.schema.children[0].contents.format == b"+r"
    u = e.import_as_table(dim=2)
    q = compile_op(rq.multidim_op())
    a, b = q(t), q(u)
    for c in range(a.ncols):
        assert np.array_equal(a.column(c), b.column(c))


@pytest.mark.parametrize("fmt", ["file", "stream"])
def test_load_arrow_file(device, tmp_path, fmt):
    """An Arrow IPC file (several record batches, nulls) loads into one
    device table holding all its rows."""
    import pyarrow as pa
    import pyarrow.ipc as ipc
    from reffine_amd import load_arrow_file
    batch, cols, valid = _batch(n=4000, seed=8)
    path = str(tmp_path / f"t.{fmt}")
    opener = ipc.new_file if fmt == "file" else ipc.new_stream
    with opener(path, batch.schema) as w:
        for off in (0, 1000, 2500):
            w.write_batch(batch.slice(off, {0: 1000, 1000: 1500, 2500: 1500}[off]))
    t = load_arrow_file(path)
    assert t.nrows == 4000
    assert np.array_equal(t.column(0), cols["a"])
    assert np.array_equal(t.validity(3), valid["d"])
