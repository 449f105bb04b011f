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


def scan_c2(v, d, q, d0=1000, vlo=0.05, vhi=0.07, qmax=24, rowid=None):
    """COUNT, SUM(q), SUM(v*q) and the ascending selection vector."""
    m = (d >= d0) & (d < d0 + 365) & (v >= vlo) & (v <= vhi) & (q < qmax)
    idx = np.nonzero(m)[0]
    sel = (rowid[idx] if rowid is not None else idx).astype(np.int32)
    return (int(m.sum()), int(q[m].sum(dtype=np.int64)),
            fsum(v[m] * q[m].astype(np.float64)), sel)
