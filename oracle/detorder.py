"""The fixed reduction order of the GPU scan kernels, restated in numpy
(test infrastructure).

Order (kernels/rf_device.h, deterministic mode):
  * rows [base, end) are cut into wave-tiles of SUB*64*ITEMS rows; in
    sub-step s lane l owns rows tile0 + (s*64 + l)*ITEMS + [0, ITEMS);
  * a lane folds its rows in ascending order from the identity:
    acc = pred ? acc (+) x : acc;
  * the wave combines lanes by xor-butterflies o = 1, 2, 4, ..., 32 with
    lane_i <- lane_i (+) lane_{i^o}; lane 0 holds the tile partial;
  * the completing block's thread t folds tile partials t, t+BLOCK, ... in
    order; each wave combines its lanes by the same butterflies and the
    block folds its waves in order w0, w1, ...;
  * the host returns init (+) total.
numpy float64 arithmetic is IEEE binary64 with round-to-nearest, so this
reproduces the kernel's result bit for bit (for + on FP64 and int64).

The same order holds for the max and min folds, with the kernels' combines
max(a, b) = b if a < b else a and min(a, b) = b if b < a else a (std::max
and std::min): not commutative once a NaN is involved, so a NaN that reaches
a lane or a partial survives or vanishes by its position in this order.
"""
from __future__ import annotations

import numpy as np


def _add(a, b):
    return a + b


def _max(a, b):
    """std::max(a, b): b if a < b else a."""
    return np.where(a < b, b, a)


def _min(a, b):
    """std::min(a, b): b if b < a else a."""
    return np.where(b < a, b, a)


OPS = {"add": _add, "max": _max, "min": _min}


def _identity(op: str, dtype):
    if op == "add":
        # -0.0 for floats: the exact additive identity the kernels fold from.
        return np.array(-0.0 if np.issubdtype(dtype, np.floating) else 0, dtype)
    if np.issubdtype(dtype, np.floating):
        return np.array(-np.inf if op == "max" else np.inf, dtype)
    info = np.iinfo(dtype)
    return np.array(info.min if op == "max" else info.max, dtype)


def _butterfly(v: np.ndarray, f=_add) -> np.ndarray:
    """v: [..., 64] lane values -> lane 0 after xor butterflies
    (lane_i <- f(lane_i, lane_{i^o}))."""
    lanes = np.arange(64)
    for o in (1, 2, 4, 8, 16, 32):
        v = f(v, v[..., lanes ^ o])
    return v[..., 0]


def _block(v: np.ndarray, block: int, f=_add) -> np.ndarray:
    """v: [..., block] thread values -> block total in the fixed order."""
    w = _butterfly(v.reshape(v.shape[:-1] + (block // 64, 64)), f)
    acc = w[..., 0]
    for i in range(1, block // 64):
        acc = f(acc, w[..., i])
    return acc


def det_fold(x: np.ndarray, pred: np.ndarray, lo: int, end: int, block: int, items: int,
             sub: int, op: str = "add", init=None):
    """Fold of x[i] over pred[i], i in [lo, end), with `op` ("add", "max",
    "min") in the kernels' fixed order. x and pred index absolute rows
    (length >= end). `init` is the query's initial state, combined last
    (init (+) total); None means the identity."""
    dtype = x.dtype
    f = OPS[op]
    ident = _identity(op, dtype)
    init = ident if init is None else np.array(init, dtype=dtype)
    base = lo - (lo % items)
    tile = sub * 64 * items
    ntiles = (end - base + tile - 1) // tile if end > lo else 0
    if ntiles == 0:
        return f(init, ident)
    n = ntiles * tile
    xs = np.full(n, ident, dtype=dtype)
    ps = np.zeros(n, dtype=bool)
    rows = np.arange(base, base + n)
    ok = (rows >= lo) & (rows < end)
    xs[ok] = x[rows[ok]]
    ps[ok] = pred[rows[ok]]
    xs = xs.reshape(ntiles, sub, 64, items)
    ps = ps.reshape(ntiles, sub, 64, items)
    acc = np.full((ntiles, 64), ident, dtype=dtype)
    for s in range(sub):
        for k in range(items):
            acc = np.where(ps[:, s, :, k], f(acc, xs[:, s, :, k]), acc)
    parts = _butterfly(acc, f)  # [ntiles]
    # completing block: thread t folds parts t, t+block, ...
    nt = np.full(block, ident, dtype=dtype)
    for i in range(0, ntiles, block):
        chunk = parts[i:i + block]
        nt[:len(chunk)] = f(nt[:len(chunk)], chunk)
    total = _block(nt[None, :], block, f)[0]
    return f(init, total)


def det_sum(x: np.ndarray, pred: np.ndarray, lo: int, end: int, block: int, items: int,
            sub: int, init=0.0):
    """Sum of x[i] over pred[i], i in [lo, end), in the kernels' fixed order.
    x and pred index absolute rows (length >= end)."""
    return det_fold(x, pred, lo, end, block, items, sub, "add", init)
