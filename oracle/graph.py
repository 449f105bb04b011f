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


def nbody(b: dict, n: int, g: float = 1.0, dt: float = 0.01):
    """Restates reffine's Nbody step (benchmark/include/misc.h:77-202):
    the pair table in the nested loops' order (a, then b; a == b skipped,
    its rows compacted), each force one IEEE operation at a time as the C
    text writes it; per body idx the sequential sum of pair rows
    [idx*n, (idx+1)*n) that exist; then the updated velocities and
    positions. Returns the 8 output columns (id, M, X, Y, Z, VX, VY, VZ)."""
    m, x, y, z = b["m"], b["x"], b["y"], b["z"]
    a_i, b_i = np.divmod(np.arange(n * n, dtype=np.int64), n)
    keep = a_i != b_i
    a_i, b_i = a_i[keep], b_i[keep]
    dx, dy, dz = x[a_i] - x[b_i], y[a_i] - y[b_i], z[a_i] - z[b_i]
    dist2 = (dx * dx + dy * dy) + dz * dz
    dist = np.sqrt(dist2)
    f = (g * (m[a_i] * m[b_i])) / dist2
    pf = [(f * dx) / dist, (f * dy) / dist, (f * dz) / dist]
    rows = len(a_i)
    force = np.zeros((3, n))
    for idx in range(n):
        lo, hi = idx * n, min((idx + 1) * n, rows)
        for c in range(3):
            s = 0.0
            for r in range(lo, hi):
                s = s + float(pf[c][r])
            force[c, idx] = s
    out_v = [b["vx"] + (force[0] / m) * dt, b["vy"] + (force[1] / m) * dt,
             b["vz"] + (force[2] / m) * dt]
    out_p = [x + out_v[0] * dt, y + out_v[1] * dt, z + out_v[2] * dt]
    return [np.arange(n, dtype=np.int64), m.copy(), *out_p, *out_v]


def pagerank_np(t: dict, n: int, alpha: float = 0.85):
    """pagerank() vectorised over runs (each run still folded in row
    order, position by position, in float32) for large graphs."""
    e, r, pr = t["edges"], t["rev_edges"], t["pr"]
    s = e["src"]
    cut = np.flatnonzero(np.diff(s)) + 1
    lo = np.concatenate([[0], cut]) if len(s) else np.zeros(0, np.int64)
    deg = np.add.reduceat(e["count"], lo) if len(s) else np.zeros(0, np.int64)
    keys = s[lo]
    contrib = np.zeros(n, dtype=np.float32)
    contrib[keys] = pr["rank"][keys] / deg.astype(np.float32)
    d = r["dst"]
    cut = np.flatnonzero(np.diff(d)) + 1
    lo = np.concatenate([[0], cut])
    hi = np.concatenate([cut, [len(d)]])
    acc = np.zeros(len(lo), dtype=np.float32)
    term = r["count"].astype(np.float32) * contrib[r["src"]]
    ln = hi - lo
    for j in range(int(ln.max()) if len(ln) else 0):
        m = ln > j
        acc[m] = acc[m] + term[lo[m] + j]
    out = np.float32(alpha) * acc + np.float32((1 - alpha) / n)
    return d[lo], out.astype(np.float32)


def nbody_np(b: dict, n: int, g: float = 1.0, dt: float = 0.01):
    """nbody() with the window sums folded position by position."""
    m, x, y, z = b["m"], b["x"], b["y"], b["z"]
    a_i, b_i = np.divmod(np.arange(n * n, dtype=np.int64), n)
    keep = a_i != b_i
    a_i, b_i = a_i[keep], b_i[keep]
    dx, dy, dz = x[a_i] - x[b_i], y[a_i] - y[b_i], z[a_i] - z[b_i]
    dist2 = (dx * dx + dy * dy) + dz * dz
    dist = np.sqrt(dist2)
    f = (g * (m[a_i] * m[b_i])) / dist2
    pf = np.stack([(f * dx) / dist, (f * dy) / dist, (f * dz) / dist])
    rows = pf.shape[1]
    force = np.zeros((3, n))
    for j in range(n):
        r = np.arange(n) * n + j
        ok = r < rows
        force[:, ok] = force[:, ok] + pf[:, r[ok]]
    out_v = [b["vx"] + (force[0] / m) * dt, b["vy"] + (force[1] / m) * dt,
             b["vz"] + (force[2] / m) * dt]
    out_p = [x + out_v[0] * dt, y + out_v[1] * dt, z + out_v[2] * dt]
    return [np.arange(n, dtype=np.int64), m.copy(), *out_p, *out_v]
