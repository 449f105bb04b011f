"""Hash join + grouped aggregate on the GPU (the nested reffine form of
reffine_amd.workloads.group_join_c4) against the oracle.

Group keys and COUNTs are bit-exact in both modes. In the deterministic
mode each group's FP64 sum is a left fold in ascending probe-row order from
0.0 -- the order of the reference's inner loop -- so it is compared
bit-exactly with np.add.at (sequential). The default mode folds each
partition of groups with LDS atomics (any order): its sums are within 1e-12
relative of the sequential fold for every group and of the correctly
rounded sum (math.fsum) for the first groups.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import brute
from reffine_amd import DeviceTable, ReffineError, compile_op, datagen, workloads

pytestmark = pytest.mark.gpu


def _upload(b, p, pvalid=None, bvalid=None):
    B = DeviceTable.from_columns([b["bk"], b["bd"], b["bf"]], dim=1, validity=bvalid)
    P = DeviceTable.from_columns([p["pk"], p["pv"], p["pq"]], dim=1, validity=pvalid)
    return B, P


def _cols(out):
    return out.column(0), out.column(1), out.column(2)


def _close(got, want, tol=1e-12):
    """Elementwise |got - want| <= tol * |want| (exact when want is 0)."""
    return bool(np.all(np.abs(got - want) <= tol * np.abs(want)))


def _check(out, b, p, bf_eq=1, pq_gt=10, fsum_groups=2000, det=False):
    keys, sums, counts, ranks, vals = brute.group_join_c4(b, p, bf_eq, pq_gt)
    k, s, c = _cols(out)
    assert np.array_equal(k, keys)
    assert np.array_equal(c, counts)
    if det:
        assert np.array_equal(s.view(np.int64), sums.view(np.int64)), "fold order differs"
    else:
        assert _close(s, sums)
    ng = min(len(keys), fsum_groups)
    if ng:
        sel = ranks < ng
        want = brute.group_fsums(ng, ranks[sel], vals[sel])
        err = max(brute.rel_err(float(a), float(w)) for a, w in zip(s[:ng], want))
        assert err <= 1e-12


@pytest.mark.parametrize("det", [False, True])
def test_group_join_matches_nested_loops(device, det):
    b, p = datagen.join_tables(300, 6000, seed=5)
    B, P = _upload(b, p)
    q = compile_op(workloads.group_join_c4(), deterministic=det)
    out = q(B, P)
    want = brute.group_join_nested(b, p)
    k, s, c = _cols(out)
    assert [int(x) for x in k] == [w[0] for w in want]
    assert [int(z) for z in c] == [w[2] for w in want]
    ws = np.array([w[1] for w in want])
    if det:
        assert [float(y) for y in s] == list(ws)  # the nested loop's order, bit for bit
    else:
        assert _close(s, ws)


# (50, 2M): ~10 groups of ~32K joined rows each -- folded a wave per group.
@pytest.mark.parametrize("nb,np_", [(1, 10), (1000, 1), (5000, 100_000), (100_000, 2_000_000),
                                    (50, 2_000_000)])
def test_group_join_sizes(device, nb, np_):
    b, p = datagen.join_tables(nb, np_, seed=nb + np_)
    B, P = _upload(b, p)
    out = compile_op(workloads.group_join_c4())(B, P)
    _check(out, b, p)


def test_group_join_full_size(device):
    """1.5M build rows, 60M probe rows (about 300K groups, 9.6M joined), in
    both modes."""
    b, p = datagen.join_tables(1_500_000, 60_000_000, seed=11)
    B, P = _upload(b, p)
    q = compile_op(workloads.group_join_c4())
    out = q(B, P)
    _check(out, b, p)
    out2 = q(B, P)  # a second run reuses the workspace
    assert np.array_equal(out2.column(0), out.column(0))
    assert np.array_equal(out2.column(2), out.column(2))
    assert _close(out2.column(1), out.column(1))
    qd = compile_op(workloads.group_join_c4(), deterministic=True)
    outd = qd(B, P)
    _check(outd, b, p, det=True)
    outd2 = qd(B, P)
    assert np.array_equal(outd2.column(1).view(np.int64), outd.column(1).view(np.int64))


@pytest.mark.parametrize("nb,np_", [(1000, 200_000), (100_000, 2_000_000)])
def test_group_join_deterministic_sizes(device, nb, np_):
    b, p = datagen.join_tables(nb, np_, seed=nb + 3 * np_)
    B, P = _upload(b, p)
    _check(compile_op(workloads.group_join_c4(), deterministic=True)(B, P), b, p, det=True)


def test_group_join_constants_and_empty(device):
    b, p = datagen.join_tables(2000, 50_000, seed=3)
    B, P = _upload(b, p)
    # no qualifying build key
    out = compile_op(workloads.group_join_c4(bf_eq=9))(B, P)
    assert out.nrows == 0
    # no qualifying probe row: every group has COUNT 0 and SUM 0.0
    out = compile_op(workloads.group_join_c4(pq_gt=100))(B, P)
    k, s, c = _cols(out)
    assert len(k) == int((b["bf"] == 1).sum()) and not c.any() and not s.any()
    for bf_eq, pq_gt in ((0, 1), (4, 49)):
        _check(compile_op(workloads.group_join_c4(bf_eq, pq_gt))(B, P), b, p, bf_eq, pq_gt)


def test_group_join_keys_not_dense(device):
    """Negative, sparse and extreme keys (the index holds any int64)."""
    rng = np.random.default_rng(4)
    nb = 3000
    keys = np.unique(rng.integers(-2**62, 2**62, size=nb, dtype=np.int64))
    keys[0] = np.iinfo(np.int64).min
    keys[-1] = np.iinfo(np.int64).max
    keys = rng.permutation(keys)
    nb = len(keys)
    b = {"bk": keys, "bd": np.zeros(nb, np.int32), "bf": rng.integers(0, 2, nb).astype(np.int8)}
    n = 200_000
    pk = keys[rng.integers(0, nb, n)]
    pk[::7] = rng.integers(-2**62, 2**62, size=len(pk[::7]), dtype=np.int64)  # mostly misses
    p = {"pk": pk, "pv": rng.random(n), "pq": rng.integers(1, 51, n).astype(np.int64)}
    B, P = _upload(b, p)
    _check(compile_op(workloads.group_join_c4())(B, P), b, p)


def _sparse_join(seed, nb=3000, n=200_000):
    rng = np.random.default_rng(seed)
    keys = rng.permutation(np.unique(rng.integers(-2**62, 2**62, size=nb, dtype=np.int64)))
    nb = len(keys)
    b = {"bk": keys, "bd": np.zeros(nb, np.int32), "bf": rng.integers(0, 2, nb).astype(np.int8)}
    pk = keys[rng.integers(0, nb, n)]
    pk[::5] = rng.integers(-2**62, 2**62, size=len(pk[::5]), dtype=np.int64)
    p = {"pk": pk, "pv": rng.random(n), "pq": rng.integers(1, 51, n).astype(np.int64)}
    return b, p


def test_group_join_index_scale(device):
    """The hash index over sparse build keys at 2, 4 and 8 slots per key
    gives the same bits in the deterministic mode (the load factor changes
    where keys sit, not which row a key finds); scales below 2 are refused
    (a full table would leave misses no empty slot to stop at)."""
    b, p = _sparse_join(6)
    B, P = _upload(b, p)
    outs = []
    for sc in (2, 4, 8):
        q = compile_op(workloads.group_join_c4(), deterministic=True, index_scale=sc)
        assert f"index scale {sc}" in q.describe()
        out = q(B, P)
        _check(out, b, p, det=True)
        outs.append([out.column(k) for k in range(3)])
    for o in outs[1:]:
        assert all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(o, outs[0]))
    with pytest.raises(ReffineError):
        compile_op(workloads.group_join_c4(), index_scale=1)


def test_group_join_autotune_picks_index_scale(device, tmp_path, monkeypatch):
    """The autotuner sweeps the index load factor of a group join at its
    best launch shape and persists both; `tuned` compiles them again."""
    import json

    from reffine_amd import autotune
    monkeypatch.setenv("RF_TUNING_CACHE", str(tmp_path / "tuning.json"))
    monkeypatch.setattr(autotune, "_mem", {})
    b, p = _sparse_join(7, nb=20_000, n=1_000_000)
    B, P = _upload(b, p)
    fn = workloads.group_join_c4()
    best, times = autotune.autotune(fn, [B, P], cands=[(256, 2, 2, 16, 0, 0)], scales=(2, 8))
    assert {k[1] for k in times} == {None, 2, 8}
    _check(best(B, P), b, p)
    e = next(iter(json.loads((tmp_path / "tuning.json").read_text()).values()))
    assert e["index_scale"] in (None, 2, 8)
    monkeypatch.setattr(autotune, "_mem", {})
    again = autotune.tuned(fn, [B, P])
    assert again.tuning() == best.tuning() and again.index_scale == e["index_scale"]
    _check(again(B, P), b, p)


def test_group_join_nulls(device):
    """Null build keys are not iterated; null probe keys join nothing
    (the validity of the iterator column, iter_space.cpp:60-66)."""
    b, p = datagen.join_tables(4000, 100_000, seed=8)
    rng = np.random.default_rng(1)
    bv = rng.random(4000) > 0.1
    pv = rng.random(100_000) > 0.2
    B, P = _upload(b, p, pvalid=[pv, None, None], bvalid=[bv, None, None])
    out = compile_op(workloads.group_join_c4())(B, P)
    b2 = {k: v[bv] for k, v in b.items()}
    p2 = {k: v[pv] for k, v in p.items()}
    _check(out, b2, p2)


# 1000 build rows: the probe pre-aggregates in LDS; 20000: the partitioned
# fold. Both builds are dense (no host synchronisation): the duplicate is
# found at the run's last synchronisation.
@pytest.mark.parametrize("nb", [1000, 20_000])
def test_group_join_duplicate_build_keys_rejected(device, nb):
    b, p = datagen.join_tables(nb, 10_000, seed=2)
    b["bk"] = b["bk"].copy()
    b["bf"] = np.ones(nb, np.int8)
    b["bk"][10] = b["bk"][20]
    B, P = _upload(b, p)
    q = compile_op(workloads.group_join_c4())
    with pytest.raises(ReffineError, match="more than once"):
        q(B, P)
    with pytest.raises(ReffineError, match="more than once"):
        q(B, P)  # the failed build is not kept
    good, _ = datagen.join_tables(nb, 10, seed=3)
    G = DeviceTable.from_columns([good["bk"], good["bd"], good["bf"]], dim=1)
    _check(q(G, P), good, p)  # the query stays usable


@pytest.mark.parametrize("nb", [2000, 20_000])
def test_group_join_nothing_qualifies(device, nb):
    """No qualifying build key (an empty output), and no qualifying probe row
    (every group COUNT 0, SUM 0.0), on both default-mode folds."""
    b, p = datagen.join_tables(nb, 100_000, seed=nb)
    B, P = _upload(b, p)
    assert compile_op(workloads.group_join_c4(bf_eq=9))(B, P).nrows == 0
    k, s, c = _cols(compile_op(workloads.group_join_c4(pq_gt=100))(B, P))
    assert np.array_equal(k, np.sort(b["bk"][b["bf"] == 1]))
    assert not c.any() and not s.any()


def test_group_join_timing(device):
    b, p = datagen.join_tables(50_000, 1_000_000, seed=9)
    B, P = _upload(b, p)
    q = compile_op(workloads.group_join_c4())
    total, scan = q.time_kernels(B, P, iters=3)
    assert total > 0 and 0 < scan <= total


def test_group_join_build_cache_is_keyed_and_clearable(device):
    """The resident build side is reused only for the same build table: runs
    alternating between two build tables, and runs after the cache is
    cleared, all give the oracle's answer."""
    b1, p = datagen.join_tables(5000, 200_000, seed=41)
    b2, _ = datagen.join_tables(5000, 10, seed=42)
    B1, P = _upload(b1, p)
    B2 = DeviceTable.from_columns([b2["bk"], b2["bd"], b2["bf"]], dim=1)
    q = compile_op(workloads.group_join_c4())
    for B, b in ((B1, b1), (B2, b2), (B1, b1)):
        _check(q(B, P), b, p)
    q.clear_cache()
    _check(q(B2, P), b2, p)
    q.clear_cache()
    q.clear_cache()
    _check(q(B1, P), b1, p)


def _five_field_join():
    """group_join_c4 with a five-field state: SUM(pv*pq) f64, COUNT, MAX(pq)
    from -1, MIN(pq) from 1000 and COUNT of the rows with pv < 0.5 (a
    predicated field)."""
    from reffine_amd.reffiner import (_cast, _eq, _f64, _f64_t, _func, _get, _gt, _i8, _i64,
                                      _i64_t, _in, _lt, _max, _min, _mul, _new, _op, _red,
                                      _sel, _sym, _add)
    B = _sym("B", workloads.C4_BUILD_TYPE)
    P = _sym("P", workloads.C4_PROBE_TYPE)
    t = _sym("t", _i64_t)
    i = _sym("i", _i64_t)
    pe = P[i]
    rows = _op([i], _in(i, P) & _eq(i, t) & _gt(pe[1], _i64(10)),
               [_mul(pe[0], _cast(_f64_t, pe[1])), pe[1], pe[0]])

    def acc(s, r):
        return _new([_add(_get(s, 0), r[1]), _add(_get(s, 1), _i64(1)), _max(_get(s, 2), r[2]),
                     _min(_get(s, 3), r[2]),
                     _sel(_lt(r[3], _f64(0.5)), _add(_get(s, 4), _i64(1)), _get(s, 4))])

    red = _red(rows, lambda: _new([_f64(0), _i64(0), _i64(-1), _i64(1000), _i64(0)]), acc)
    grp = _sym("grp", red)
    be = B[t]
    out = _op([t], _in(t, B) & _eq(be[1], _i8(1)),
              [_get(grp, 0), _get(grp, 1), _get(grp, 2), _get(grp, 3), _get(grp, 4)])
    os_ = _sym("out", out)
    fn = _func("five_field_join", os_, [B, P])
    fn.tbl[grp] = red
    fn.tbl[os_] = out
    return fn


def _five_field_oracle(b, p):
    keys = np.sort(np.asarray(b["bk"])[np.asarray(b["bf"]) == 1])
    pk, pv, pq = (np.asarray(p[c]) for c in ("pk", "pv", "pq"))
    m = pq > 10
    pos = np.searchsorted(keys, pk[m])
    ok = (pos < len(keys)) & (keys[np.minimum(pos, max(len(keys) - 1, 0))] == pk[m])
    g, v, q = pos[ok], pv[m][ok], pq[m][ok]
    sums = np.zeros(len(keys))
    np.add.at(sums, g, v * q.astype(np.float64))
    counts = np.bincount(g, minlength=len(keys)).astype(np.int64)
    mx = np.full(len(keys), -1, dtype=np.int64)
    np.maximum.at(mx, g, q)
    mn = np.full(len(keys), 1000, dtype=np.int64)
    np.minimum.at(mn, g, q)
    low = np.bincount(g[v < 0.5], minlength=len(keys)).astype(np.int64)
    return keys, sums, counts, mx, mn, low


# 300 and 1000 build rows: few groups (five 8-byte slots per group within
# 48 KiB of LDS per block), the probe scan pre-aggregates in LDS (default
# mode); 20000: the partitioned fold. Deterministic: sort + fold.
@pytest.mark.parametrize("nb", [300, 1000, 20_000])
@pytest.mark.parametrize("det", [False, True])
def test_group_join_five_fields(device, nb, det):
    b, p = datagen.join_tables(nb, 400_000, seed=nb + 17)
    B, P = _upload(b, p)
    q = compile_op(_five_field_join(), deterministic=det)
    want = _five_field_oracle(b, p)
    for _ in range(2):  # the second run: the build side from the cache
        out = q(B, P)
        got = [out.column(j) for j in range(6)]
        assert np.array_equal(got[0], want[0])
        for j in (2, 3, 4, 5):
            assert np.array_equal(got[j], want[j]), j
        if det:
            assert np.array_equal(got[1].view(np.int64), want[1].view(np.int64))
        else:
            assert _close(got[1], want[1])


def test_group_join_lds_preaggregation_empty_probe(device):
    """Few groups and no joined row (no probe row passes pq > 10): every
    qualifying build key gets the reduction's init, as its inner loop over
    zero rows gives."""
    b, p = datagen.join_tables(500, 10_000, seed=29)
    p["pq"] = np.minimum(p["pq"], 10)
    B, P = _upload(b, p)
    out = compile_op(_five_field_join())(B, P)
    keys = np.sort(np.asarray(b["bk"])[np.asarray(b["bf"]) == 1])
    assert np.array_equal(out.column(0), keys)
    assert np.all(out.column(1) == 0.0) and np.all(out.column(2) == 0)
    assert np.all(out.column(3) == -1) and np.all(out.column(4) == 1000)
    assert np.all(out.column(5) == 0)
