"""Restatement of reffine's TPC-H Q3/Q4/Q12/Q18 over a run-end lineitem
(test infrastructure: tests and the CPU baseline only).

The reference runs each query as a merge of two sorted key sets
(`_in(k, lineitem) & _in(k, orders)`: InterSpace, src/iter/iter_space.cpp:
318-364) that visits the keys present in both in ascending order; for each
key, `lineitem[k]` is the key's run of rows (run ends of column 0,
src/pass/loopgen.cpp:34-38) folded row by row from the init
(loopgen.cpp:206-269), `orders[k]` is the key's orders row, and
`_locate(customer, x)` is the row of x in customer's index or -1
(include/reffine/vinstr/vinstr.h:137-146). The query bodies follow
benchmark/include/tpch.h line by line (cited per function). Plain Python
loops over the rows: every FP64 operation is one IEEE double operation in
the order the C text performs it.
"""
from __future__ import annotations

import numpy as np


def _runs(keys: np.ndarray):
    """(key, lo, hi) per run of equal adjacent keys, in storage order."""
    if len(keys) == 0:
        return []
    cut = np.flatnonzero(np.diff(keys)) + 1
    lo = np.concatenate([[0], cut])
    hi = np.concatenate([cut, [len(keys)]])
    return [(int(keys[a]), int(a), int(b)) for a, b in zip(lo, hi)]


def _merged(li: dict, od: dict):
    """Keys in both lineitem runs and orders, ascending, with the run's row
    range and the orders row."""
    orow = {int(k): i for i, k in enumerate(od["o_orderkey"].tolist())}
    out = []
    for k, lo, hi in _runs(li["l_orderkey"]):
        if k in orow:
            out.append((k, lo, hi, orow[k]))
    out.sort()
    return out


def q3(t: dict, segment: int = 1, date: int = 795484800):
    """tpch.h:30-73. Returns (orderkeys, revenue)."""
    li, od, cu = t["lineitem"], t["orders"], t["customer"]
    price, disc = li["l_extendedprice"].tolist(), li["l_discount"].tolist()
    ship = li["l_shipdate"].tolist()
    crow = {int(k): i for i, k in enumerate(cu["c_custkey"].tolist())}
    seg = cu["c_mktsegment"].tolist()
    keys, vals = [], []
    for k, lo, hi, o in _merged(li, od):
        s = 0.0
        for r in range(lo, hi):
            if ship[r] > date:
                s = s + price[r] * (1.0 - disc[r])
        c = crow.get(int(od["o_custkey"][o]), -1)
        if (c >= 0 and int(od["o_orderdate"][o]) < date and int(seg[c]) == segment
                and s > 0.0):
            keys.append(k)
            vals.append(s)
    return np.array(keys, dtype=np.int64), np.array(vals, dtype=np.float64)


def q4(t: dict, start_date: int = 700000000, end_date: int = 900000000):
    """tpch.h:98-148. Returns the five int32 counts by order priority."""
    li, od = t["lineitem"], t["orders"]
    comm, rec = li["l_commitdate"].tolist(), li["l_receiptdate"].tolist()
    cnt = [0] * 5
    for k, lo, hi, o in _merged(li, od):
        exists = False
        for r in range(lo, hi):
            exists = (comm[r] < rec[r]) or exists
        d = int(od["o_orderdate"][o])
        if d >= start_date and d < end_date and exists:
            p = int(od["o_orderpriority"][o])
            for j in range(5):
                if p == j:
                    cnt[j] = np.int32(cnt[j] + 1)
    return [int(c) for c in cnt]


def q12(t: dict):
    """tpch.h:545-604. Returns (orderkeys, high counts, low counts)."""
    li, od = t["lineitem"], t["orders"]
    comm, ship, rec = (li["l_commitdate"].tolist(), li["l_shipdate"].tolist(),
                       li["l_receiptdate"].tolist())
    keys, hi_c, lo_c = [], [], []
    for k, lo, hi, o in _merged(li, od):
        p = int(od["o_orderpriority"][o])
        cond = p == 0 or p == 1
        hc, lc = (1 if cond else 0), (0 if cond else 1)
        s0 = s1 = 0
        for r in range(lo, hi):
            if comm[r] < rec[r] and ship[r] < comm[r]:
                s0, s1 = hc + s0, lc + s1
        keys.append(k)
        hi_c.append(s0)
        lo_c.append(s1)
    return (np.array(keys, dtype=np.int64), np.array(hi_c, dtype=np.int32),
            np.array(lo_c, dtype=np.int32))


def q18(t: dict, quantity: int = 300):
    """tpch.h:684-715. Returns (orderkeys, total quantity)."""
    li, od = t["lineitem"], t["orders"]
    qty = li["l_quantity"].tolist()
    keys, vals = [], []
    for k, lo, hi, o in _merged(li, od):
        s = 0.0
        for r in range(lo, hi):
            s = s + qty[r]
        if s > float(quantity):
            keys.append(k)
            vals.append(s)
    return np.array(keys, dtype=np.int64), np.array(vals, dtype=np.float64)


# ---------------------------------------------------------------------------
# The same four queries vectorised over runs (numpy), for tables too large
# for the row loops above. Each run is still folded in row order: position
# j of every run at once, j = 0, 1, ... (a sequential fold per run).

def _runs_np(keys: np.ndarray):
    n = len(keys)
    if n == 0:
        z = np.zeros(0, dtype=np.int64)
        return z, z, z
    cut = np.flatnonzero(np.diff(keys)) + 1
    lo = np.concatenate([[0], cut]).astype(np.int64)
    hi = np.concatenate([cut, [n]]).astype(np.int64)
    return keys[lo], lo, hi


def _fold_np(lo, hi, vals, op, init):
    """Per run, init op vals[lo] op vals[lo+1] ... in row order."""
    s = np.full(len(lo), init, dtype=np.asarray(vals).dtype if init is None else type(init))
    ln = hi - lo
    for j in range(int(ln.max()) if len(ln) else 0):
        m = ln > j
        s[m] = op(s[m], vals[lo[m] + j])
    return s


def _match_np(li, od):
    """Runs of lineitem whose key is in orders, ascending: (run keys, run
    lo, run hi, orders row)."""
    rk, lo, hi = _runs_np(li["l_orderkey"])
    ok = od["o_orderkey"]
    order = np.argsort(ok, kind="stable")
    pos = np.searchsorted(ok[order], rk)
    pos = np.minimum(pos, max(len(ok) - 1, 0))
    hit = (len(ok) > 0) & (ok[order][pos] == rk) if len(ok) else np.zeros(len(rk), bool)
    orow = order[pos]
    sel = np.flatnonzero(hit)
    srt = sel[np.argsort(rk[sel], kind="stable")]
    return rk[srt], lo[srt], hi[srt], orow[srt]


def q18_np(t: dict, quantity: int = 300):
    k, lo, hi, _ = _match_np(t["lineitem"], t["orders"])
    s = _fold_np(lo, hi, t["lineitem"]["l_quantity"], np.add, 0.0)
    m = s > float(quantity)
    return k[m], s[m]


def q3_np(t: dict, segment: int = 1, date: int = 795484800):
    li, od, cu = t["lineitem"], t["orders"], t["customer"]
    k, lo, hi, o = _match_np(li, od)
    rev = li["l_extendedprice"] * (1.0 - li["l_discount"])
    s = np.zeros(len(lo))
    ln = hi - lo
    for j in range(int(ln.max()) if len(ln) else 0):
        m = ln > j
        r = lo[m] + j
        s[m] = np.where(li["l_shipdate"][r] > date, s[m] + rev[r], s[m])
    ck = cu["c_custkey"]
    corder = np.argsort(ck, kind="stable")
    cust = od["o_custkey"][o]
    cp = np.minimum(np.searchsorted(ck[corder], cust), len(ck) - 1)
    found = ck[corder][cp] == cust
    seg = cu["c_mktsegment"][corder[cp]]
    m = found & (od["o_orderdate"][o] < date) & (seg == segment) & (s > 0.0)
    return k[m], s[m]


def q4_np(t: dict, start_date: int = 700000000, end_date: int = 900000000):
    li, od = t["lineitem"], t["orders"]
    k, lo, hi, o = _match_np(li, od)
    late = li["l_commitdate"] < li["l_receiptdate"]
    ex = _fold_np(lo, hi, late, np.logical_or, False)
    d = od["o_orderdate"][o]
    m = (d >= start_date) & (d < end_date) & ex
    p = od["o_orderpriority"][o][m]
    return [int((p == j).sum()) for j in range(5)]


def q12_np(t: dict):
    li, od = t["lineitem"], t["orders"]
    k, lo, hi, o = _match_np(li, od)
    ok = ((li["l_commitdate"] < li["l_receiptdate"]) &
          (li["l_shipdate"] < li["l_commitdate"])).astype(np.int32)
    n = _fold_np(lo, hi, ok, np.add, np.int32(0))
    p = od["o_orderpriority"][o]
    high = (p == 0) | (p == 1)
    return k, np.where(high, n, 0).astype(np.int32), np.where(high, 0, n).astype(np.int32)


def example(t: dict, nation_key: int = 0, size: int = 15):
    """tpch.h:606-682. Per partsupp run (a part's suppliers, in storage
    order), the availqty of rows whose supplier's nation is nation_key,
    folded in row order as int32; parts of size > `size`. Returns
    (partkeys, quantities)."""
    ps, part, sup = t["partsupp"], t["part"], t["supplier"]
    prow = {int(k): i for i, k in enumerate(part["p_partkey"].tolist())}
    srow = {int(k): i for i, k in enumerate(sup["s_suppkey"].tolist())}
    nation = sup["s_nationkey"].tolist()
    skey, qty = ps["ps_suppkey"].tolist(), ps["ps_availqty"].tolist()
    psize = part["p_size"].tolist()
    keys, vals = [], []
    for k, lo, hi in _runs(ps["ps_partkey"]):
        v = np.int32(0)
        for r in range(lo, hi):
            if int(nation[srow[skey[r]]]) == nation_key:
                v = np.int32(v + np.int32(qty[r]))
        if int(psize[prow[k]]) > size:
            keys.append(k)
            vals.append(int(v))
    return np.array(keys, dtype=np.int64), np.array(vals, dtype=np.int32)


def example_np(t: dict, nation_key: int = 0, size: int = 15):
    ps, part, sup = t["partsupp"], t["part"], t["supplier"]
    rk, lo, hi = _runs_np(ps["ps_partkey"])
    sorder = np.argsort(sup["s_suppkey"], kind="stable")
    si = sorder[np.searchsorted(sup["s_suppkey"][sorder], ps["ps_suppkey"])]
    contrib = np.where(sup["s_nationkey"][si] == nation_key, ps["ps_availqty"], 0).astype(np.int32)
    v = _fold_np(lo, hi, contrib, np.add, np.int32(0))
    porder = np.argsort(part["p_partkey"], kind="stable")
    pi = porder[np.searchsorted(part["p_partkey"][porder], rk)]
    m = part["p_size"][pi] > size
    return rk[m], v[m]


# ---------------------------------------------------------------------------
# The SubVector folds of tpch.h: Q1, Q6 and TPC-DS Q9. Counts exact; float64
# sums correctly rounded (math.fsum), the value the GPU's tree order is
# checked against within 1e-12.

def _fsum(x) -> float:
    import math
    return math.fsum(np.asarray(x, dtype=np.float64).tolist()) if len(x) else 0.0


def q1(t: dict):
    """tpch.h:398-486: rows [0, number of lineitem runs) (_len of the
    run-end column 0), three l_returnflag groups of SUM(qty), SUM(price),
    SUM(price*(1-disc)), SUM(price*(1-disc)*(1+tax)), COUNT."""
    li = t["lineitem"]
    n = len(_runs_np(li["l_orderkey"])[0])
    rf = li["l_returnflag"][:n]
    qty, ep = li["l_quantity"][:n], li["l_extendedprice"][:n]
    disc, tax = li["l_discount"][:n], li["l_tax"][:n]
    disc_price = ep * (1.0 - disc)
    charge = disc_price * (1.0 + tax)
    out = []
    for g in range(3):
        m = rf == g
        out += [_fsum(qty[m]), _fsum(ep[m]), _fsum(disc_price[m]), _fsum(charge[m]),
                float(m.sum())]
    return out


def q6(t: dict, start: int = 820454400, end: int = 852076800, disc: float = 0.05,
       quant: float = 24.5):
    """tpch.h:160-216: SUM(price * discount) over the qualifying lines."""
    li = t["lineitem"]
    d, q = li["l_discount"], li["l_quantity"]
    m = ((li["l_shipdate"] >= start) & (li["l_shipdate"] < end) & (d >= disc - 0.01) &
         (d <= disc + 0.01) & (q < quant))
    return _fsum(li["l_extendedprice"][m] * d[m])


def tpcds_q9(ss: dict):
    """tpch.h:218-313: five quantity buckets of (COUNT, SUM(ext_tax),
    SUM(net_paid_inc_tax))."""
    q = ss["ss_quantity"]
    out = []
    for b in range(5):
        m = (q >= 20 * b + 1) & (q <= 20 * b + 20)
        out.append((int(m.sum()), _fsum(ss["ss_ext_tax"][m]), _fsum(ss["ss_net_paid_inc_tax"][m])))
    return out


def q11(t: dict, nation_key: int = 0):
    """tpch.h:315-396 (the function's output, op): per partsupp run, the
    sum of availqty * supplycost over rows whose supplier is in
    nation_key, folded in row order; parts with a positive sum."""
    ps, sup = t["partsupp"], t["supplier"]
    srow = {int(k): i for i, k in enumerate(sup["s_suppkey"].tolist())}
    nation = sup["s_nationkey"].tolist()
    skey, qty = ps["ps_suppkey"].tolist(), ps["ps_availqty"].tolist()
    cost = ps["ps_supplycost"].tolist()
    keys, vals = [], []
    for k, lo, hi in _runs(ps["ps_partkey"]):
        s = 0.0
        for r in range(lo, hi):
            if int(nation[srow[skey[r]]]) == nation_key:
                s = s + float(qty[r]) * cost[r]
        if s > 0.0:
            keys.append(k)
            vals.append(s)
    return np.array(keys, dtype=np.int64), np.array(vals, dtype=np.float64)
