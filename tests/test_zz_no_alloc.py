"""The no-allocation entry point (rf_query_run_noalloc) against rf_query_run
on the same tables: the same reduction struct bit for bit and the same
selection rows, on two calls in a row (the second over a smaller table), in
both reduction modes; its limits come back as status codes before any
launch (or, for a selection larger than out_cap, after placing out_cap
rows)."""
import ctypes
import struct

import numpy as np
import pytest

from reffine_amd import DeviceTable, compile_op, datagen, workloads
from reffine_amd.runtime import lib

pytestmark = pytest.mark.gpu

RF_OK, RF_ERR_INVALID, RF_ERR_NOMEM = 0, 1, 4


def c2_table(n, seed, validity=None):
    c = datagen.scan_table(n, seed, with_q=True, rowid_dtype=np.int32)
    cols = [c["rowid"], c["k"], c["v"], c["d"], c["q"]]
    return DeviceTable.from_columns(cols, dim=1, validity=validity)


def c1_table(n, seed):
    c = datagen.scan_table(n, seed, with_q=False, rowid_dtype=np.int64)
    return DeviceTable.from_columns([c["rowid"], c["k"], c["v"], c["d"]], dim=2)


def run_noalloc(q, inputs, out=None, cap=0):
    arr = q._args(inputs)
    buf = ctypes.create_string_buffer(max(8, q.out_bytes))
    cols = None
    if out is not None:
        cols = (ctypes.c_void_p * out.ncols)(*[out.device_ptr(c) for c in range(out.ncols)])
    rows = ctypes.c_int64(-1)
    st = lib().rf_query_run_noalloc(q.handle, arr, len(inputs), buf, cols, cap,
                                    ctypes.byref(rows))
    return st, buf.raw[:q.out_bytes], rows.value


@pytest.mark.parametrize("det", [False, True])
def test_selection_and_sums_match_rf_query_run(device, det):
    q = compile_op(workloads.scan_select_c2(d0=400, vlo=0.0, vhi=0.3, qmax=30),
                   deterministic=det)
    cap = 400_000
    out = DeviceTable.from_columns([np.zeros(cap, np.int32)], dim=1)
    for n, seed in ((1_000_000, 3), (200_003, 4)):  # the second call is smaller
        t = c2_table(n, seed)
        (cnt, sq, svq), sel = q(t)
        st, raw, rows = run_noalloc(q, [t], out, cap)
        assert st == RF_OK, lib().rf_last_error()
        # {i64 count, i64 sum(q), f64 sum(v*q)} then the selection's handle
        # (NULL here: its rows are in the caller's buffer).
        got = struct.unpack("<qqdQ", raw)
        assert raw[:24] == struct.pack("<qqd", cnt, sq, svq)  # bit for bit
        assert got[3] == 0
        assert rows == cnt == sel.nrows and 0 < rows <= cap
        np.testing.assert_array_equal(out.column(0)[:rows], sel.column(0))


def test_reduction_only_query(device):
    q = compile_op(workloads.scan_agg_c1())
    for n, seed in ((1 << 20, 5), (5000, 6)):
        t = c1_table(n, seed)
        st, raw, rows = run_noalloc(q, [t])
        assert st == RF_OK, lib().rf_last_error()
        assert rows == 0
        assert raw == q.run_raw(t)[0]


def test_limits_are_status_codes(device):
    q = compile_op(workloads.scan_select_c2(d0=400, vlo=0.0, vhi=0.3, qmax=30))
    t = c2_table(300_000, 7)
    (cnt, _, _), sel = q(t)
    # A selection larger than out_cap: the first out_cap rows are placed.
    small = 1000
    out = DeviceTable.from_columns([np.zeros(small, np.int32)], dim=1)
    st, _, rows = run_noalloc(q, [t], out, small)
    assert st == RF_ERR_NOMEM and rows == cnt > small
    np.testing.assert_array_equal(out.column(0), sel.column(0)[:small])
    # A selection without output buffers.
    st, _, _ = run_noalloc(q, [t], None, 0)
    assert st == RF_ERR_INVALID
    # More scratch than the fixed array holds (20M rows of staging).
    big = c2_table(20_000_000, 8)
    st, _, _ = run_noalloc(q, [big], out, small)
    assert st == RF_ERR_NOMEM and b"scratch" in lib().rf_last_error()
    # A null layout whose kernel was never loaded; loaded by one plain run.
    valid = [np.arange(50_000) % 7 != 0, None, None, None, None]
    tn = c2_table(50_000, 9, validity=valid)
    out2 = DeviceTable.from_columns([np.zeros(50_000, np.int32)], dim=1)
    st, _, _ = run_noalloc(q, [tn], out2, 50_000)
    assert st == RF_ERR_INVALID and b"null layout" in lib().rf_last_error()
    (cnt_n, _, _), sel_n = q(tn)
    st, _, rows = run_noalloc(q, [tn], out2, 50_000)
    assert st == RF_OK and rows == cnt_n
    np.testing.assert_array_equal(out2.column(0)[:rows], sel_n.column(0))


def test_probe_query_over_both_index_forms(device):
    """A selection that probes a second table: its kernel variant follows
    the probed table's index form (a packed rank map over ascending keys,
    or the same keys shuffled, with a row table). Each form runs once
    through rf_query_run, then the no-allocation call gives the same rows;
    the second call is over a smaller scanned table."""
    from test_scan_gpu import _guarded_join
    rng = np.random.default_rng(5)
    m = 50_000
    i = np.arange(m, dtype=np.int64)
    bk = 100 + 3 * i + (i % 2)
    bv = rng.integers(-1000, 1000, m, dtype=np.int64)
    q = compile_op(_guarded_join(0))
    cap = 200_000
    out = DeviceTable.from_columns([np.zeros(cap, np.int64), np.zeros(cap, np.int64)], dim=1)
    for order, na in ((np.arange(m), 160_000), (rng.permutation(m), 120_001)):
        B = DeviceTable.from_columns([bk[order], bv[order]], dim=1)
        ak = np.arange(0, na, dtype=np.int64)
        av = rng.integers(-1000, 1000, na, dtype=np.int64)
        A = DeviceTable.from_columns([ak, av], dim=1)
        want = q(A, B)
        st, raw, rows = run_noalloc(q, [A, B], out, cap)
        assert st == RF_OK, lib().rf_last_error()
        assert rows == want.nrows > 10_000
        for c in range(2):
            np.testing.assert_array_equal(out.column(c)[:rows], want.column(c))
