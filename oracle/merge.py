"""The union merge (an Op over _in(t, A) | _in(t, B), the reference's
UnionSpace) restated on the CPU (test infrastructure).

The reference lowers such an Op to one loop (src/pass/loopgen.cpp:44-110)
over UnionSpace(VecSpace(A), VecSpace(B)) (src/iter/iter_space.cpp:55-91,
216-316):

    lb  = min(A.key[0], B.key[0])                  # UnionSpace::_lower_bound
    idx = (locate(A, lb), locate(B, lb))           # JointSpace::_iter_to_idx
    while idx.a < len(A) or idx.b < len(B):        # UnionSpace::_is_alive
        t = min(A.key[idx.a], B.key[idx.b])        # JointSpace::_idx_to_iter
        if (idx.a < len(A) and valid(A.key, idx.a)) or \
           (idx.b < len(B) and valid(B.key, idx.b)):   # UnionSpace::_iter_cond
            emit t, outputs(A row idx.a, B row idx.b)  # Element reads the current rows
        lc = A.key[idx.a] <= B.key[idx.b]; rc = B.key[idx.b] <= A.key[idx.a]
        idx = (idx.a + lc, idx.b + rc)             # JointSpace::_next

Where the reference reads outside a vector the loop has no defined result:
locate() returns -1 for the side that does not hold lb (when the two first
keys differ) and the loop then reads row -1, and once one side has run out
the loop reads its row one past the end. This library defines those cases
as a plain outer merge: both sides start at row 0 (what locate gives when the
first keys are equal), and a side past its last row is absent -- its key is
never the minimum and never advances, its columns read 0 and its key is not
valid. For strictly ascending keys every state the reference reaches with
both rows inside their vectors is reached here with the same outputs.

For strictly ascending keys the emitting states have a closed form: one per
key k of the sorted union of the two key sets, at A row = #(A keys < k) and
B row = #(B keys < k). `union_states_np` uses it at full sizes; it is pinned
against the loop by brute force over all small inputs.

A side written _in(t + c, X), c a constant, is a ShiftedSpace
(src/pass/reffinepass.cpp:67-71): its iterator at row i is X.key[i] + c
(ShiftedSpace::_idx_to_iter, src/iter/iter_space.cpp:136-140), so the loop
above runs on the keys plus c (`off_a`, `off_b`; wrapping int64 addition).
AlgoTrading (benchmark/include/misc.h:15-61) merges a table with itself
shifted by -50.
"""
from __future__ import annotations

import numpy as np


def _shift(keys, off):
    with np.errstate(over="ignore"):
        return np.asarray(keys, np.int64) + np.int64(off)


def union_states_loop(ka, kb, valid_a=None, valid_b=None, off_a=0, off_b=0):
    """The merge loop above, row by row: (t, i, j) of every emitting state in
    loop order (i, j: the current rows of A and B, past the end when out)."""
    ka, kb = _shift(ka, off_a), _shift(kb, off_b)
    na, nb = len(ka), len(kb)
    va = np.ones(na, bool) if valid_a is None else np.asarray(valid_a, bool)
    vb = np.ones(nb, bool) if valid_b is None else np.asarray(valid_b, bool)
    out = []
    i = j = 0
    while i < na or j < nb:
        a_in, b_in = i < na, j < nb
        x = int(ka[i]) if a_in else None
        y = int(kb[j]) if b_in else None
        if a_in and b_in:
            t = min(x, y)
        else:
            t = x if a_in else y
        if (a_in and va[i]) or (b_in and vb[j]):
            out.append((t, i, j))
        lc = a_in and (not b_in or x <= y)
        rc = b_in and (not a_in or y <= x)
        i += lc
        j += rc
    return out


def union_states_np(ka, kb, valid_a=None, valid_b=None, off_a=0, off_b=0):
    """The same states for strictly ascending keys, vectorised:
    (t, i, j) arrays in loop order."""
    ka = _shift(ka, off_a)
    kb = _shift(kb, off_b)
    t = np.union1d(ka, kb)
    i = np.searchsorted(ka, t, "left")
    j = np.searchsorted(kb, t, "left")
    na, nb = len(ka), len(kb)
    keep = np.zeros(len(t), bool)
    ina, inb = i < na, j < nb
    if valid_a is None:
        keep |= ina
    else:
        keep[ina] |= np.asarray(valid_a, bool)[i[ina]]
    if valid_b is None:
        keep |= inb
    else:
        keep[inb] |= np.asarray(valid_b, bool)[j[inb]]
    return t[keep], i[keep], j[keep]


def read_at(col, rows):
    """Column values at `rows`, 0 past the column's end (an absent side)."""
    col = np.asarray(col)
    rows = np.asarray(rows, np.int64)
    out = np.zeros(len(rows), col.dtype)
    ok = rows < len(col)
    out[ok] = col[rows[ok]]
    return out


def ojoin(ka, va, kb, vb, valid_a=None, valid_b=None):
    """benchmark/include/micro.h:79-95's outer join: (t, left[t].0 -
    right[t].0) per emitting state, as int64 (wrapping)."""
    t, i, j = union_states_np(ka, kb, valid_a, valid_b)
    with np.errstate(over="ignore"):
        return t, read_at(va, i) - read_at(vb, j)


def algotrading(ka, va, kb, vb, shift=-50):
    """benchmark/include/misc.h:15-61's AlgoTrading: an Op over
    _in(t, A) | _in(t + shift, B) with outputs (t, A[t].0 - B[t + shift].0),
    each side read at its current row; values float32 (the stock_price
    table's `val`), the difference rounded to float32."""
    t, i, j = union_states_np(ka, kb, off_b=shift)
    a = read_at(np.asarray(va, np.float32), i)
    b = read_at(np.asarray(vb, np.float32), j)
    return t, (a - b).astype(np.float32)
