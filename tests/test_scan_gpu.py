"""GPU parity of the fused scan-filter-aggregate (+ selection) kernels
against the brute-force oracle and the fixed-order restatement.

Bars: COUNT / int64 SUM / selection vectors bit-exact; FP64 sums within
1e-12 relative error of the correctly rounded sum (math.fsum); the
deterministic mode bit-exact against oracle.detorder."""
import numpy as np
import pytest

from oracle import brute, cjit, detorder
from reffine_amd import DeviceTable, compile_op, datagen, workloads

FP_TOL = 1e-12
pytestmark = pytest.mark.gpu


def c1_table(n, seed=42):
    c = datagen.scan_table(n, seed, with_q=False, rowid_dtype=np.int64)
    return c, DeviceTable.from_columns([c["rowid"], c["k"], c["v"], c["d"]], dim=2)


def c2_table(n, seed=43):
    c = datagen.scan_table(n, seed, with_q=True, rowid_dtype=np.int32)
    return c, DeviceTable.from_columns([c["rowid"], c["k"], c["v"], c["d"], c["q"]], dim=1)


@pytest.mark.parametrize("n", [0, 1, 63, 1000, 4097, 1 << 20])
def test_c1_scan_agg(device, n):
    c, t = c1_table(n)
    q = compile_op(workloads.scan_agg_c1())
    cnt, s = q(t)
    want_cnt, want_s = brute.scan_c1(c["k"], c["v"], c["d"])
    assert cnt == want_cnt
    assert brute.rel_err(s, want_s) <= FP_TOL


@pytest.mark.parametrize("n", [1, 5000, 1 << 20])
def test_c1_deterministic_bit_exact(device, n):
    c, t = c1_table(n, seed=5)
    q = compile_op(workloads.scan_agg_c1(), deterministic=True)
    cnt, s = q(t)
    m = (c["d"] >= 365) & (c["d"] < 730) & (c["v"] < 0.5)
    x = c["v"] * c["k"].astype(np.float64)
    want = detorder.det_sum(x, m, 0, n, *q.tuning()[:3])
    assert cnt == int(m.sum())
    assert s == float(want)  # bit-exact
    # and repeatable
    assert q(t) == (cnt, s)


def _bits(vals):
    return [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in vals]


@pytest.mark.parametrize("n", [4097, 1 << 20])
def test_deterministic_mode_equals_cjit_fixed_order(device, n):
    """The deterministic mode's C1 and C2 results equal the CPU JIT
    restatement's fixed-order emission (oracle/cjit.py, C text in the
    kernels' tile / butterfly / completion order) at the kernel's launch
    shape, bit for bit; C2's selection too."""
    c, t = c1_table(n, seed=5)
    q = compile_op(workloads.scan_agg_c1(), deterministic=True)
    (want,) = cjit.compile_op(workloads.scan_agg_c1(), fixed_order=q.tuning()[:3])(
        {"cols": [c["rowid"], c["k"], c["v"], c["d"]]})
    assert _bits(q(t)) == _bits(want)
    c2, t2 = c2_table(n)
    q2 = compile_op(workloads.scan_select_c2(), deterministic=True)
    st, sel = q2(t2)
    wst, wsel = cjit.compile_op(workloads.scan_select_c2(), fixed_order=q2.tuning()[:3])(
        {"cols": [c2["rowid"], c2["k"], c2["v"], c2["d"], c2["q"]]})
    assert _bits(st) == _bits(wst)
    np.testing.assert_array_equal(sel.column(0), wsel[0])


@pytest.mark.parametrize("n", [0, 17, 4096, 4097, 200_003])
def test_c2_scan_select(device, n):
    c, t = c2_table(n)
    q = compile_op(workloads.scan_select_c2())
    (cnt, sq, svq), sel = q(t)
    w_cnt, w_sq, w_svq, w_sel = brute.scan_c2(c["v"], c["d"], c["q"], rowid=c["rowid"])
    assert cnt == w_cnt and sq == w_sq
    assert brute.rel_err(svq, w_svq) <= FP_TOL
    assert sel.nrows == w_cnt
    got = sel.column(0) if sel.nrows else np.zeros(0, np.int32)
    assert got.dtype == np.int32
    np.testing.assert_array_equal(got, w_sel)


def test_c2_full_size_60m(device):
    n = 60_000_000
    c, t = c2_table(n, seed=2)
    q = compile_op(workloads.scan_select_c2())
    (cnt, sq, svq), sel = q(t)
    w_cnt, w_sq, w_svq, w_sel = brute.scan_c2(c["v"], c["d"], c["q"], rowid=c["rowid"])
    assert cnt == w_cnt and sq == w_sq
    assert brute.rel_err(svq, w_svq) <= FP_TOL
    np.testing.assert_array_equal(sel.column(0), w_sel)
    # Second run on the same plan (reused workspace, new epoch) is identical.
    (cnt2, sq2, svq2), sel2 = q(t)
    assert (cnt2, sq2) == (cnt, sq) and brute.rel_err(svq2, w_svq) <= FP_TOL
    np.testing.assert_array_equal(sel2.column(0), w_sel)


def test_c2_with_nulls(device):
    n = 100_000
    c = datagen.scan_table(n, 9, with_q=True, rowid_dtype=np.int32)
    valid0 = datagen.uniform(9, 99, n) > 0.1
    t = DeviceTable.from_columns([c["rowid"], c["k"], c["v"], c["d"], c["q"]], dim=1,
                                 validity=[valid0, None, None, None, None])
    q = compile_op(workloads.scan_select_c2())
    (cnt, sq, svq), sel = q(t)
    m = ((c["d"] >= 1000) & (c["d"] < 1365) & (c["v"] >= 0.05) & (c["v"] <= 0.07) &
         (c["q"] < 24) & valid0)
    assert cnt == int(m.sum()) and sq == int(c["q"][m].sum())
    assert brute.rel_err(svq, brute.fsum(c["v"][m] * c["q"][m])) <= FP_TOL
    np.testing.assert_array_equal(sel.column(0), c["rowid"][m])


def test_autotune_keeps_results(device, tmp_path, monkeypatch):
    """The autotuner's choice (a few candidates on a 2M-row C2 table)
    computes the same count, integer sum and selection as the default shape;
    it is persisted per query shape and `tuned` compiles it again from the
    file (another table of a similar row count shares it)."""
    import json

    from reffine_amd import autotune
    monkeypatch.setenv("RF_TUNING_CACHE", str(tmp_path / "tuning.json"))
    monkeypatch.setattr(autotune, "_mem", {})
    n = 2_000_000
    c = datagen.scan_table(n, 31, with_q=True, rowid_dtype=np.int32)
    t = DeviceTable.from_columns([c["rowid"], c["k"], c["v"], c["d"], c["q"]], dim=1)
    fn = workloads.scan_select_c2()
    cands = [(256, 2, 4, 16, 0, 0), (256, 2, 16, 8, 1, 0), (256, 4, 8, 8, 1, 4)]
    best, times = autotune.autotune(fn, [t], cands=cands)
    assert {k[0] for k in times} == set(cands) and all(ms > 0 for ms in times.values())
    (cnt, sq, _), sel = best(t)
    (cnt0, sq0, _), sel0 = compile_op(fn)(t)
    assert (cnt, sq) == (cnt0, sq0)
    assert np.array_equal(sel.column(0), sel0.column(0))
    saved = json.loads((tmp_path / "tuning.json").read_text())
    assert len(saved) == 1 and tuple(next(iter(saved.values()))["tuning"]) == best.tuning()
    monkeypatch.setattr(autotune, "_mem", {})  # from the file
    c2 = datagen.scan_table(n + 12345, 32, with_q=True, rowid_dtype=np.int32)
    t2 = DeviceTable.from_columns([c2["rowid"], c2["k"], c2["v"], c2["d"], c2["q"]], dim=1)
    again = autotune.tuned(fn, [t2])
    assert again.tuning() == best.tuning()


def _max_query(det_init: float, use_min: bool = False):
    """MAX (or MIN) of x over rows with d < 100, state first in the fold
    (std::max(s, x)), over T(rowid i64, x f64, d i32), dim 2."""
    from reffine_amd.reffiner import (_f64, _f64_t, _func, _get, _i32, _i32_t, _i64_t, _idx,
                                      _len, _lt, _max, _min, _red, _sel, _subvec, _sym, _vec_t)
    T = _sym("T", _vec_t(2, _i64_t, _f64_t, _i32_t))
    f = _min if use_min else _max

    def acc(s, v):
        return _sel(_lt(_get(v, 1), _i32(100)), f(s, _get(v, 0)), s)

    red = _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _f64(det_init), acc)
    rs = _sym("m", red)
    fn = _func("maxfold", rs, [T])
    fn.tbl[rs] = red
    return fn


@pytest.mark.parametrize("use_min", [False, True])
def test_deterministic_minmax_nan_and_signed_zero(device, use_min):
    """Float MAX / MIN folds in the deterministic mode against the fixed
    order (butterfly levels as exact xor pairs): rows hold NaNs and both
    signs of zero, on which std::max / std::min are not commutative."""
    n = 300_000
    rng = np.random.default_rng(11)
    x = np.where(rng.random(n) < 0.5, 0.0, -0.0)
    x = np.where(rng.random(n) < 0.05, np.nan, x)
    x = np.where(rng.random(n) < 0.3, (-1.0 if not use_min else 1.0) * rng.random(n), x)
    d = rng.integers(0, 200, n).astype(np.int32)
    t = DeviceTable.from_columns([np.arange(n, dtype=np.int64), x, d], dim=2)
    init = -5.0 if not use_min else 5.0
    for det in (True, False):
        q = compile_op(_max_query(init, use_min), deterministic=det)
        got = q(t)
        m = d < 100
        want_val = (np.max if not use_min else np.min)(np.where(np.isnan(x[m]), -np.inf if not use_min else np.inf, x[m]))
        assert got == want_val  # value (either sign of zero)
        if det:
            want = detorder.det_fold(x, m, 0, n, *q.tuning()[:3], op="min" if use_min else "max",
                                     init=init)
            assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)


def _guarded_join(thr):
    """_op over t in A with A[t].v > thr and t in B: (t, A[t].v + B[t].v).
    The probe-free conjunct A[t].v > thr guards the probe of B's index
    (joins through _in as in test/src/test_join.cpp:9-35)."""
    from reffine_amd.reffiner import _add, _func, _gt, _i64, _i64_t, _in, _op, _sym, _vec_t
    ty = _vec_t(1, _i64_t, _i64_t)
    a, b = _sym("avec", ty), _sym("bvec", ty)
    t = _sym("t", _i64_t)
    ae, be = a[t][0], b[t][0]
    a_s, b_s = _sym("a", ae), _sym("b", be)
    op = _op([t], _in(t, a) & _gt(a_s, _i64(thr)) & _in(t, b), [_add(a_s, b_s)])
    op_sym = _sym("op", op)
    fn = _func("guarded_join", op_sym, [a, b])
    fn.tbl[a_s] = ae
    fn.tbl[b_s] = be
    fn.tbl[op_sym] = op
    return fn


@pytest.mark.parametrize("ascending", [True, False])
def test_guarded_probe_misses(device, ascending):
    """A scan's probe of a rank map: B's keys strictly ascending with gaps
    (probed without branches) or the same rows shuffled (a row table
    behind the map). A's keys fall below B's range, past it and into its
    interior gaps, and the guard A[t].v > thr is false on about half the
    rows; the output equals the numpy join."""
    rng = np.random.default_rng(11)
    m = 300_000
    i = np.arange(m, dtype=np.int64)
    bk = 100 + 3 * i + (i % 2)
    bv = rng.integers(-1000, 1000, m, dtype=np.int64)
    ak = np.arange(0, int(bk[-1]) + 200, dtype=np.int64)
    av = rng.integers(-1000, 1000, len(ak), dtype=np.int64)
    order = np.arange(m) if ascending else rng.permutation(m)
    A = DeviceTable.from_columns([ak, av], dim=1)
    B = DeviceTable.from_columns([bk[order], bv[order]], dim=1)
    q = compile_op(_guarded_join(0))
    out = q(A, B)
    pos = np.searchsorted(bk, ak)
    hit = (pos < m) & (bk[np.minimum(pos, m - 1)] == ak)
    keep = hit & (av > 0)
    assert keep.sum() > 100_000 and (hit & ~(av > 0)).sum() > 100_000
    np.testing.assert_array_equal(out.column(0), ak[keep])
    np.testing.assert_array_equal(out.column(1), av[keep] + bv[np.minimum(pos, m - 1)][keep])
    # A second table of the other index form through the same query.
    B2 = DeviceTable.from_columns([bk[::-1].copy(), bv[::-1].copy()], dim=1) if ascending \
        else DeviceTable.from_columns([bk, bv], dim=1)
    out2 = q(A, B2)
    np.testing.assert_array_equal(out2.column(0), ak[keep])
    np.testing.assert_array_equal(out2.column(1), av[keep] + bv[np.minimum(pos, m - 1)][keep])
