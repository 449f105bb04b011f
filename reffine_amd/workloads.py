"""Query builders for the scan workloads, written in reffine's own idioms.

The scan-filter-aggregate form is reffine's TPC-H Q6 idiom
(benchmark/include/tpch.h:174-208): `_red(_subvec(T, 0, _len(T, 1)), init,
acc)` with `acc = _sel(pred, new_state, s)` over a 2-dimensional table whose
fold skips column 0. The selection form is an `_op` over `_in(t, T)`
(test/src/test_transform.cpp:9-28) whose iterator column is an int32 row
id, so its output column 0 is the int32 selection vector.
"""
from __future__ import annotations

from .reffiner import (_add, _cast, _eq, _f32_t, _f64, _f64_t, _func, _gt, _gte, _i8, _i8_t, _i32, _i32_t, _i64,
                       _i64_t, _sub, _idx, _in, _len, _lt, _lte, _mul, _new, _op, _red, _sel, _subvec,
                       _sym, _vec_t, _get)


def scan_agg_c1(lo: int = 365, hi: int = 730, vmax: float = 0.5):
    """COUNT(*), SUM(v*k) WHERE lo <= d < hi AND v < vmax over
    T(rowid i64, k i64, v f64, d i32), dim 2."""
    T = _sym("T", _vec_t(2, _i64_t, _i64_t, _f64_t, _i32_t))

    def acc(s, v):
        k, x, d = _get(v, 0), _get(v, 1), _get(v, 2)
        pred = _gte(d, _i32(lo)) & _lt(d, _i32(hi)) & _lt(x, _f64(vmax))
        new_s = _new([_add(_get(s, 0), _i64(1)),
                      _add(_get(s, 1), _mul(x, _cast(_f64_t, k)))])
        return _sel(pred, new_s, s)

    red = _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _new([_i64(0), _f64(0)]), acc)
    rs = _sym("red", red)
    fn = _func("scan_c1", rs, [T])
    fn.tbl[rs] = red
    return fn


C2_TYPE = _vec_t(1, _i32_t, _i64_t, _f64_t, _i32_t, _i64_t)  # rowid, k, v, d, q


def _c2_pred(T, t, d0, vlo, vhi, qmax):
    e = T[t]
    v, d, q = e[1], e[2], e[3]
    return (_in(t, T) & _gte(d, _i32(d0)) & _lt(d, _i32(d0 + 365)) & _gte(v, _f64(vlo)) &
            _lte(v, _f64(vhi)) & _lt(q, _i64(qmax))), e


def scan_select_c2(d0: int = 1000, vlo: float = 0.05, vhi: float = 0.07, qmax: int = 24,
                   with_selection: bool = True):
    """COUNT, SUM(q) i64, SUM(v*q) f64 and the int32 selection vector of
    rows with d in [d0, d0+365), vlo <= v <= vhi, q < qmax, over
    T(rowid i32, k i64, v f64, d i32, q i64). Output struct
    {{i64, i64, f64}, selection-vector} (the selection is an Op whose only
    column is the iterator)."""
    T = _sym("T", C2_TYPE)
    t = _sym("t", _i32_t)
    pred, e = _c2_pred(T, t, d0, vlo, vhi, qmax)
    rows = _op([t], pred, [e[3], e[1]])  # iterator, q, v

    def acc(s, r):
        q, v = r[1], r[2]
        return _new([_add(_get(s, 0), _i64(1)), _add(_get(s, 1), q),
                     _add(_get(s, 2), _mul(v, _cast(_f64_t, q)))])

    red = _red(rows, lambda: _new([_i64(0), _i64(0), _f64(0)]), acc)
    rs = _sym("agg", red)
    if not with_selection:
        fn = _func("scan_c2_agg", rs, [T])
        fn.tbl[rs] = red
        return fn
    t2 = _sym("t", _i32_t)
    pred2, _ = _c2_pred(T, t2, d0, vlo, vhi, qmax)
    sel = _op([t2], pred2, [])
    ss = _sym("sel", sel)
    out = _new([rs, ss])
    os_ = _sym("out", out)
    fn = _func("scan_c2", os_, [T])
    fn.tbl[rs] = red
    fn.tbl[ss] = sel
    fn.tbl[os_] = out
    return fn


def c2_bytes_per_row() -> int:
    """Algorithmic bytes per scanned row: d (4) + v (8) + q (8)."""
    return 20


# rowid, f1, f2, qty, price, disc, tax, d (2-dim: the fold reads columns 1..)
C3_TYPE = _vec_t(2, _i64_t, _i8_t, _i8_t, _i64_t, _f64_t, _f64_t, _f64_t, _i32_t)
C3_GROUPS = [(a, b) for a in range(3) for b in range(2)]
C3_FIELDS = ["sum_qty", "sum_price", "sum_disc_price", "sum_charge", "sum_disc", "count"]


def group_agg_c3(dmax: int | None = 2526):
    """Low-cardinality grouped aggregate in the TPC-H Q1 idiom
    (benchmark/include/tpch.h:412-478): one unrolled
    `_sel(f1 == a & f2 == b & d <= dmax, s + x, s)` per (group, field).
    State: for each of the 6 groups (f1, f2) in order (0,0), (0,1), ...,
    (2,1): SUM(qty) i64, SUM(price), SUM(price*(1-disc)),
    SUM(price*(1-disc)*(1+tax)), SUM(disc) f64, COUNT i64. AVG(qty),
    AVG(price), AVG(disc) are the sums over COUNT. With dmax=None the bound
    is a scalar i32 input of the function (a prepared query)."""
    T = _sym("T", C3_TYPE)
    dsym = _sym("dmax", _i32_t) if dmax is None else None

    def acc(s, v):
        f1, f2, qty, price, disc, tax, d = (_get(v, i) for i in range(7))
        disc_price = _mul(price, _sub(_f64(1), disc))
        charge = _mul(disc_price, _add(_f64(1), tax))
        keep = _lte(d, dsym if dsym is not None else _i32(dmax))
        vals = [qty, price, disc_price, charge, disc, _i64(1)]
        out = []
        for g, (a, b) in enumerate(C3_GROUPS):
            p = keep & _eq(f1, _i8(a)) & _eq(f2, _i8(b))
            for k, x in enumerate(vals):
                sk = _get(s, g * len(vals) + k)
                out.append(_sel(p, _add(sk, x), sk))
        return _new(out)

    def init():
        return _new([_i64(0) if k in (0, 5) else _f64(0)
                     for _ in C3_GROUPS for k in range(len(C3_FIELDS))])

    red = _red(_subvec(T, _idx(0), _len(T, 1)), init, acc)
    rs = _sym("q1", red)
    fn = _func("group_c3", rs, [T] if dsym is None else [T, dsym])
    fn.tbl[rs] = red
    return fn


def c3_bytes_per_row() -> int:
    """f1 (1) + f2 (1) + qty (8) + price (8) + disc (8) + tax (8) + d (4)."""
    return 38


# Hash join + high-cardinality grouped aggregate.
C4_BUILD_TYPE = _vec_t(1, _i64_t, _i32_t, _i8_t)   # bk, bd, bf
C4_PROBE_TYPE = _vec_t(1, _i64_t, _f64_t, _i64_t)  # pk, pv, pq


def group_join_c4(bf_eq: int = 1, pq_gt: int = 10):
    """SUM(pv*pq) f64 and COUNT i64 per build key bk with bf == bf_eq, over
    the probe rows with pk == bk and pq > pq_gt; one output row per
    qualifying build key, in key order.

    Written as reffine's nested form: an Op over the build table whose
    outputs reduce an Op over the probe table correlated by `i == t`
    (nested Ops as in test/src/test_nested.cpp:9-25, joins through `_in`
    as in test/src/test_join.cpp:9-44, the reduction as in
    test/src/test_reduce.cpp:61-122). Output columns: bk, SUM, COUNT.
    """
    B = _sym("B", C4_BUILD_TYPE)
    P = _sym("P", C4_PROBE_TYPE)
    t = _sym("t", _i64_t)
    i = _sym("i", _i64_t)
    pe = P[i]
    rows = _op([i], _in(i, P) & _eq(i, t) & _gt(pe[1], _i64(pq_gt)),
               [_mul(pe[0], _cast(_f64_t, pe[1]))])

    def acc(s, r):
        return _new([_add(_get(s, 0), r[1]), _add(_get(s, 1), _i64(1))])

    red = _red(rows, lambda: _new([_f64(0), _i64(0)]), acc)
    grp = _sym("grp", red)
    be = B[t]
    out = _op([t], _in(t, B) & _eq(be[1], _i8(bf_eq)), [_get(grp, 0), _get(grp, 1)])
    os_ = _sym("out", out)
    fn = _func("group_join_c4", os_, [B, P])
    fn.tbl[grp] = red
    fn.tbl[os_] = out
    return fn


def c4_bytes(n_build: int, n_probe: int, groups: int) -> int:
    """Algorithmic bytes: build bk + bf (9 B/row), probe pk + pv + pq
    (24 B/row), output bk + SUM + COUNT (24 B/group)."""
    return 9 * n_build + 24 * n_probe + 24 * groups


# Two keyed vectors for the union merge (micro.h's fake_data: key, value).
MERGE_TYPE = _vec_t(1, _i64_t, _i64_t)


def outer_join(vt=MERGE_TYPE):
    """benchmark/include/micro.h:79-95's outer join: an Op over
    `_in(t, left) | _in(t, right)` (UnionSpace) whose outputs are t and
    left[t].0 - right[t].0, each side read at its current row of the merge."""
    left = _sym("left", vt)
    right = _sym("right", vt)
    t = _sym("t", _i64_t)
    diff = _sub(_get(left[t], 0), _get(right[t], 0))
    dsym = _sym("diff", diff)
    op = _op([t], _in(t, left) | _in(t, right), [dsym])
    osym = _sym("op", op)
    fn = _func("ojoinbench", osym, [left, right])
    fn.tbl[osym] = op
    fn.tbl[dsym] = diff
    return fn


# benchmark/include/misc.h's stock_price table: ts (int64, ascending), val (float32).
STOCK_TYPE = _vec_t(1, _i64_t, _f32_t)


def algotrading(shift: int = -50):
    """benchmark/include/misc.h:15-61's AlgoTrading: an Op over
    `_in(t, stock_price1) | _in(t + shift, stock_price2)` (a union with a
    shifted side) whose output is stock_price1[t].0 - stock_price2[t + shift].0,
    each side read at its current row of the merge."""
    sp1 = _sym("stock_price1", STOCK_TYPE)
    sp2 = _sym("stock_price2", STOCK_TYPE)
    t = _sym("t", _i64_t)
    t50 = _add(t, _i64(shift))
    diff = _sub(_get(sp1[t], 0), _get(sp2[t50], 0))
    dsym = _sym("diff", diff)
    op = _op([t], _in(t, sp1) | _in(t50, sp2), [dsym])
    osym = _sym("op", op)
    fn = _func("algotrading", osym, [sp1, sp2])
    fn.tbl[osym] = op
    fn.tbl[dsym] = diff
    return fn


def merge_bytes(na: int, nb: int, n_out: int) -> int:
    """Algorithmic bytes of the outer join: both sides' keys and values
    (16 B/row) read, (t, diff) written (16 B per output row)."""
    return 16 * (na + nb) + 16 * n_out


# The grouped layout iterated by row id (an Op over `_in(t, T)` needs a
# one-dimensional vector; the same device columns serve both views).
C3_DIM1_TYPE = _vec_t(1, _i64_t, _i8_t, _i8_t, _i64_t, _f64_t, _f64_t, _f64_t, _i32_t)


def q6_template():
    """Scan-filter-aggregate + selection over the grouped layout (the
    TPC-H Q6 shape of the 60M-row scan): WHERE d0 <= d < d0+365 AND
    dlo <= disc <= dhi AND qty < qmax -> COUNT, SUM(qty) i64,
    SUM(price*disc) f64 and the selected row ids. The predicate constants
    are scalar inputs (d0 i32, dlo f64, dhi f64, qmax i64), so one compiled
    query serves every constant choice."""
    T = _sym("T", C3_DIM1_TYPE)
    d0 = _sym("d0", _i32_t)
    dlo = _sym("dlo", _f64_t)
    dhi = _sym("dhi", _f64_t)
    qmax = _sym("qmax", _i64_t)

    def pred(t):
        e = T[t]
        qty, disc, d = e[2], e[4], e[6]
        return (_in(t, T) & _gte(d, d0) & _lt(d, _add(d0, _i32(365))) & _gte(disc, dlo) &
                _lte(disc, dhi) & _lt(qty, qmax)), e

    t = _sym("t", _i64_t)
    p, e = pred(t)
    rows = _op([t], p, [e[2], e[3], e[4]])  # iterator, qty, price, disc

    def acc(s, r):
        return _new([_add(_get(s, 0), _i64(1)), _add(_get(s, 1), r[1]),
                     _add(_get(s, 2), _mul(r[2], r[3]))])

    red = _red(rows, lambda: _new([_i64(0), _i64(0), _f64(0)]), acc)
    rs = _sym("agg", red)
    t2 = _sym("t", _i64_t)
    p2, _ = pred(t2)
    sel = _op([t2], p2, [])
    ss = _sym("sel", sel)
    out = _new([rs, ss])
    os_ = _sym("out", out)
    fn = _func("q6_batch", os_, [T, d0, dlo, dhi, qmax])
    fn.tbl[rs] = red
    fn.tbl[ss] = sel
    fn.tbl[os_] = out
    return fn


def c5_mix(n_queries: int, n_batches: int, seed: int = 17, templates=("q6", "q1", "join")):
    """The request mix: per query a template (uniform), a target record
    batch (uniform) and seeded predicate constants with at most 6 decimals.
    Returns a list of (template, batch, params)."""
    from . import datagen
    import numpy as np
    tpl = datagen.randint(seed, 1, n_queries, 0, len(templates), np.int64)
    bat = datagen.randint(seed, 2, n_queries, 0, n_batches, np.int64)
    d0 = datagen.randint(seed, 3, n_queries, 0, 2192, np.int64)
    lo = datagen.randint(seed, 4, n_queries, 0, 9, np.int64)
    qm = datagen.randint(seed, 5, n_queries, 10, 41, np.int64)
    dm = datagen.randint(seed, 6, n_queries, 2000, 2556, np.int64)
    fl = datagen.randint(seed, 7, n_queries, 0, 5, np.int64)
    out = []
    for i in range(n_queries):
        name = templates[int(tpl[i])]
        if name == "q6":
            params = (int(d0[i]), round(int(lo[i]) * 0.01, 2), round(int(lo[i]) * 0.01 + 0.02, 2),
                      int(qm[i]))
        elif name == "q1":
            params = (int(dm[i]),)
        else:
            params = (int(fl[i]), int(qm[i]))
        out.append((name, int(bat[i]), params))
    return out


# Date dimension joined with the grouped layout's d column.
C5_DIM_TYPE = _vec_t(1, _i64_t, _i8_t)  # dk (day number), dflag


def dim_join_template():
    """Hash join of a record batch (grouped layout) with the date dimension
    on d = dk, for the days with dflag == f: per day, SUM(price*(1-disc)) f64
    and COUNT i64 over the rows with qty < qmax; one row per qualifying day,
    in key order. The nested form of group_join_c4 with the build and probe
    constants as scalar inputs (f i8, qmax i64)."""
    D = _sym("D", C5_DIM_TYPE)
    T = _sym("T", C3_DIM1_TYPE)
    f = _sym("f", _i8_t)
    qmax = _sym("qmax", _i64_t)
    t = _sym("t", _i64_t)
    i = _sym("i", _i64_t)
    e = T[i]
    rows = _op([i], _in(i, T) & _eq(_cast(_i64_t, e[6]), t) & _lt(e[2], qmax),
               [_mul(e[3], _sub(_f64(1), e[4]))])

    def acc(s, r):
        return _new([_add(_get(s, 0), r[1]), _add(_get(s, 1), _i64(1))])

    red = _red(rows, lambda: _new([_f64(0), _i64(0)]), acc)
    grp = _sym("grp", red)
    de = D[t]
    out = _op([t], _in(t, D) & _eq(de[0], f), [_get(grp, 0), _get(grp, 1)])
    os_ = _sym("out", out)
    fn = _func("dim_join", os_, [D, T, f, qmax])
    fn.tbl[grp] = red
    fn.tbl[os_] = out
    return fn
