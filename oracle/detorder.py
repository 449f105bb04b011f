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
"""
from __future__ import annotations

import numpy as np


def _butterfly(v: np.ndarray) -> np.ndarray:
    """v: [..., 64] lane values -> lane 0 after xor butterflies."""
    lanes = np.arange(64)
    for o in (1, 2, 4, 8, 16, 32):
        v = v + v[..., lanes ^ o]
    return v[..., 0]


def _block(v: np.ndarray, block: int) -> np.ndarray:
    """v: [..., block] thread values -> block total in the fixed order."""
    w = _butterfly(v.reshape(v.shape[:-1] + (block // 64, 64)))
    acc = w[..., 0]
    for i in range(1, block // 64):
        acc = acc + w[..., i]
    return acc


def det_sum(x: np.ndarray, pred: np.ndarray, lo: int, end: int, block: int, items: int,
            sub: int, init=0.0):
    """Sum of x[i] over pred[i], i in [lo, end), in the kernels' fixed order.
    x and pred index absolute rows (length >= end)."""
    dtype = x.dtype
    base = lo - (lo % items)
    tile = sub * 64 * items
    ntiles = (end - base + tile - 1) // tile if end > lo else 0
    if ntiles == 0:
        return np.array(init, dtype=dtype) + np.zeros((), dtype)
    n = ntiles * tile
    xs = np.zeros(n, dtype=dtype)
    ps = np.zeros(n, dtype=bool)
    rows = np.arange(base, base + n)
    ok = (rows >= lo) & (rows < end)
    xs[ok] = x[rows[ok]]
    ps[ok] = pred[rows[ok]]
    xs = xs.reshape(ntiles, sub, 64, items)
    ps = ps.reshape(ntiles, sub, 64, items)
    acc = np.zeros((ntiles, 64), dtype=dtype)
    for s in range(sub):
        for k in range(items):
            acc = np.where(ps[:, s, :, k], acc + xs[:, s, :, k], acc)
    parts = _butterfly(acc)  # [ntiles]
    # completing block: thread t folds parts t, t+block, ...
    nt = np.zeros(block, dtype=dtype)
    for i in range(0, ntiles, block):
        chunk = parts[i:i + block]
        nt[:len(chunk)] = nt[:len(chunk)] + chunk
    total = _block(nt[None, :], block)[0]
    return np.array(init, dtype=dtype) + total
