"""Restatement of reffine's PageRank step (benchmark/include/misc.h:212-291)
-- test infrastructure only.

The reference folds each node's out-edges run into its degree, divides its
rank by it (float32), indexes the result by node, and folds each node's
in-edges run in row order: s = s + (float)count * contrib[src], then
alpha * s + (1 - alpha) / n -- every operation one float32 operation.
"""
from __future__ import annotations

import numpy as np


def _runs(keys):
    if len(keys) == 0:
        return []
    cut = np.flatnonzero(np.diff(keys)) + 1
    lo = np.concatenate([[0], cut])
    hi = np.concatenate([cut, [len(keys)]])
    return [(int(keys[a]), int(a), int(b)) for a, b in zip(lo, hi)]


def pagerank(t: dict, n: int, alpha: float = 0.85):
    """Returns (dst keys, new ranks float32) in rev_edges run order."""
    e, r, pr = t["edges"], t["rev_edges"], t["pr"]
    rank = {int(i): np.float32(x) for i, x in zip(pr["id"].tolist(), pr["rank"].tolist())}
    contrib = {}
    cnt = e["count"].tolist()
    for k, lo, hi in _runs(e["src"]):
        if k not in rank:
            continue
        deg = 0
        for i in range(lo, hi):
            deg += cnt[i]
        contrib[k] = np.float32(rank[k] / np.float32(deg))
    a = np.float32(alpha)
    c0 = np.float32((1 - alpha) / n)
    rsrc, rcnt = r["src"].tolist(), r["count"].tolist()
    keys, vals = [], []
    for k, lo, hi in _runs(r["dst"]):
        s = np.float32(0)
        for i in range(lo, hi):
            s = np.float32(s + np.float32(np.float32(rcnt[i]) * contrib.get(rsrc[i], np.float32(0))))
        keys.append(k)
        vals.append(np.float32(np.float32(a * s) + c0))
    return np.array(keys, dtype=np.int64), np.array(vals, dtype=np.float32)
