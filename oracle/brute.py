"""Brute-force restatement of the scan-path queries (test infrastructure).

Each function follows the row semantics of reffine's generated loop
(src/pass/loopgen.cpp:206-269: a row folds into the state iff its predicate
holds) and of its Op compaction (loopgen.cpp:113-204: qualifying rows are
written in iteration order). Integer results are exact; FP64 sums are
correctly rounded (math.fsum, Shewchuk's exact summation), which is the
compensated host oracle the 1e-12 tolerance is measured against. Per-row
FP64 products are computed with the same single IEEE multiply the C text
`v * (double)q` performs.
"""
from __future__ import annotations

import math

import numpy as np


def fsum(x: np.ndarray) -> float:
    return math.fsum(np.asarray(x, dtype=np.float64).tolist()) if len(x) else 0.0


def rel_err(got: float, want: float) -> float:
    if want == 0.0:
        return abs(got)
    return abs(got - want) / abs(want)


def scan_c1(k, v, d, lo=365, hi=730, vmax=0.5):
    """COUNT(*), SUM(v*k) WHERE lo <= d < hi AND v < vmax
    (benchmark/include/tpch.h:174-208 idiom)."""
    m = (d >= lo) & (d < hi) & (v < vmax)
    return int(m.sum()), fsum(v[m] * k[m].astype(np.float64))


def group_c3_contribs(c):
    """Per-row values of the grouped aggregate, each one IEEE operation as in
    the C text: qty, price, price*(1-disc), (price*(1-disc))*(1+tax), disc."""
    disc_price = c["price"] * (1.0 - c["disc"])
    charge = disc_price * (1.0 + c["tax"])
    return [c["qty"], c["price"], disc_price, charge, c["disc"]]


def group_c3(c, dmax=2526, groups=((0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1))):
    """Per group (f1, f2): SUM(qty) i64, SUM(price), SUM(price*(1-disc)),
    SUM(price*(1-disc)*(1+tax)), SUM(disc) f64, COUNT i64 over d <= dmax
    (benchmark/include/tpch.h:412-478 idiom); flat list, group-major."""
    keep = c["d"] <= dmax
    vals = group_c3_contribs(c)
    out = []
    for a, b in groups:
        m = keep & (c["f1"] == a) & (c["f2"] == b)
        out.append(int(vals[0][m].sum(dtype=np.int64)))
        for x in vals[1:]:
            out.append(fsum(x[m]))
        out.append(int(m.sum()))
    return out


def scan_c2(v, d, q, d0=1000, vlo=0.05, vhi=0.07, qmax=24, rowid=None):
    """COUNT, SUM(q), SUM(v*q) and the ascending selection vector."""
    m = (d >= d0) & (d < d0 + 365) & (v >= vlo) & (v <= vhi) & (q < qmax)
    idx = np.nonzero(m)[0]
    sel = (rowid[idx] if rowid is not None else idx).astype(np.int32)
    return (int(m.sum()), int(q[m].sum(dtype=np.int64)),
            fsum(v[m] * q[m].astype(np.float64)), sel)


def q6_batch(c, d0, dlo, dhi, qmax):
    """The Q6-shaped batch template (reffine_amd.workloads.q6_template) over
    a grouped-layout batch: COUNT, SUM(qty) i64 and SUM(price*disc) f64 of
    the rows with d0 <= d < d0+365, dlo <= disc <= dhi, qty < qmax, and their
    ascending row ids."""
    d, disc, qty = c["d"], c["disc"], c["qty"]
    m = (d >= d0) & (d < d0 + 365) & (disc >= dlo) & (disc <= dhi) & (qty < qmax)
    return (int(m.sum()), int(qty[m].sum(dtype=np.int64)), fsum(c["price"][m] * disc[m]),
            np.asarray(c["rowid"])[m].astype(np.int64))


def group_join_c4(build, probe, bf_eq=1, pq_gt=10):
    """Hash join + grouped aggregate, brute force. For every build key with
    bf == bf_eq, in ascending key order: SUM(pv*pq) over probe rows with
    pk == key and pq > pq_gt, folded from 0.0 in ascending probe-row order
    (np.add.at applies its updates one index at a time, in order), and the
    COUNT of those rows. Returns (keys, sums, counts, ranks, values) where
    ranks/values are the joined rows in probe-row order."""
    bk = np.asarray(build["bk"], dtype=np.int64)
    keys = np.sort(bk[np.asarray(build["bf"]) == bf_eq])
    pk = np.asarray(probe["pk"], dtype=np.int64)
    pq = np.asarray(probe["pq"], dtype=np.int64)
    pv = np.asarray(probe["pv"], dtype=np.float64)
    m = pq > pq_gt
    if len(keys) == 0:
        e = np.zeros(0)
        return keys, e, e.astype(np.int64), e.astype(np.int64), e
    pos = np.searchsorted(keys, pk[m])
    ok = (pos < len(keys)) & (keys[np.minimum(pos, len(keys) - 1)] == pk[m])
    ranks = pos[ok]
    vals = (pv[m] * pq[m].astype(np.float64))[ok]
    sums = np.zeros(len(keys), dtype=np.float64)
    np.add.at(sums, ranks, vals)
    counts = np.bincount(ranks, minlength=len(keys)).astype(np.int64)
    return keys, sums, counts, ranks, vals


def group_fsums(n_groups, ranks, vals):
    """Correctly rounded per-group sums (math.fsum per group)."""
    order = np.argsort(ranks, kind="stable")
    r, v = ranks[order], vals[order]
    bounds = np.searchsorted(r, np.arange(n_groups + 1))
    return np.array([math.fsum(v[bounds[g]:bounds[g + 1]].tolist()) for g in range(n_groups)])


def group_join_nested(build, probe, bf_eq=1, pq_gt=10):
    """The nested loops of the reffine form, in pure Python (small inputs
    only): for each build row in row order whose bf == bf_eq, fold the probe
    rows with pk == bk and pq > pq_gt, in row order, into (0.0, 0) with
    s0 = s0 + pv*pq, s1 = s1 + 1. Returns the rows sorted by key."""
    out = []
    bk, bf = list(map(int, build["bk"])), list(map(int, build["bf"]))
    pk, pq = list(map(int, probe["pk"])), list(map(int, probe["pq"]))
    pv = [float(x) for x in probe["pv"]]
    for r in range(len(bk)):
        if bf[r] != bf_eq:
            continue
        t = bk[r]
        s0, s1 = 0.0, 0
        for i in range(len(pk)):
            if pk[i] == t and pq[i] > pq_gt:
                s0 = s0 + pv[i] * float(pq[i])
                s1 = s1 + 1
        out.append((t, s0, s1))
    return sorted(out)


def dim_join(dim, c, f, qmax):
    """The date-dimension join: for each day dk with dflag == f (ascending),
    SUM(price*(1-disc)) folded from 0.0 in row order and COUNT over the rows
    with d == dk and qty < qmax."""
    keys = np.sort(np.asarray(dim["dk"])[np.asarray(dim["dflag"]) == f])
    m = np.asarray(c["qty"]) < qmax
    d = np.asarray(c["d"]).astype(np.int64)[m]
    vals = (np.asarray(c["price"]) * (1.0 - np.asarray(c["disc"])))[m]
    if len(keys) == 0:
        return keys, np.zeros(0), np.zeros(0, dtype=np.int64)
    pos = np.searchsorted(keys, d)
    ok = (pos < len(keys)) & (keys[np.minimum(pos, len(keys) - 1)] == d)
    sums = np.zeros(len(keys))
    np.add.at(sums, pos[ok], vals[ok])
    counts = np.bincount(pos[ok], minlength=len(keys)).astype(np.int64)
    return keys, sums, counts
