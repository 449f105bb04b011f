"""The fixed reduction order of the GPU scan kernels, restated in numpy
(test infrastructure).

Order (kernels/rf_device.h, deterministic mode):
  * rows [base, end) are cut into tiles of SUB*BLOCK*ITEMS rows; in
    sub-step s thread t owns rows tile0 + (s*BLOCK + t)*ITEMS + [0, ITEMS);
  * a thread folds its rows in ascending order from the identity:
    acc = pred ? acc (+) x : acc;
  * a wave combines lanes by xor-butterflies o = 1, 2, 4, ..., 32 with
    lane_i <- lane_i (+) lane_{i^o}; waves are folded in order w0, w1, ...;
  * the completing block's thread t folds tile partials t, t+BLOCK, ... in
    order, then the same wave and block combine gives the device total;
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
    tile = sub * block * items
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
    xs = xs.reshape(ntiles, sub, block, items)
    ps = ps.reshape(ntiles, sub, block, items)
    acc = np.zeros((ntiles, block), dtype=dtype)
    for s in range(sub):
        for k in range(items):
            acc = np.where(ps[:, s, :, k], acc + xs[:, s, :, k], acc)
    parts = _block(acc, block)  # [ntiles]
    # completing block: thread t folds parts t, t+block, ...
    nt = np.zeros(block, dtype=dtype)
    for i in range(0, ntiles, block):
        chunk = parts[i:i + block]
        nt[:len(chunk)] = nt[:len(chunk)] + chunk
    total = _block(nt[None, :], block)[0]
    return np.array(init, dtype=dtype) + total
