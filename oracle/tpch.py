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
