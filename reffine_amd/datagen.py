"""Seeded synthetic tables for the scan / group-by / join workloads.

Every value is a pure function of (seed, column, row): a SplitMix64
finaliser over a counter, so any row range can be regenerated alone and
the same table is produced on every machine. Uniform doubles are
(x >> 11) * 2^-53; uniform integers in [lo, hi) are lo + (x >> 11) % (hi-lo).
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_C = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, col: int, n: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _G + np.uint64(col) * _C
        rows = np.arange(start, start + n, dtype=np.uint64)
        return mix64(rows + base)


def uniform(seed: int, col: int, n: int, start: int = 0) -> np.ndarray:
    return (u64(seed, col, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def randint(seed: int, col: int, n: int, lo: int, hi: int, dtype, start: int = 0) -> np.ndarray:
    x = u64(seed, col, n, start) >> np.uint64(11)
    return (np.int64(lo) + (x % np.uint64(hi - lo)).astype(np.int64)).astype(dtype)


def scan_table(n: int, seed: int = 42, with_q: bool = True, rowid_dtype=np.int64):
    """Scan workload columns: rowid, k:i64 U[0,1000), v:f64 U[0,1),
    d:i32 U[0,2556) and (optionally) q:i64 U[1,50]."""
    cols = {
        "rowid": np.arange(n, dtype=rowid_dtype),
        "k": randint(seed, 1, n, 0, 1000, np.int64),
        "v": uniform(seed, 2, n),
        "d": randint(seed, 3, n, 0, 2556, np.int32),
    }
    if with_q:
        cols["q"] = randint(seed, 4, n, 1, 51, np.int64)
    return cols


def group_table(n: int, seed: int = 7):
    """Low-cardinality group-by workload: f1 in {0,1,2} (25/25/50),
    f2 in {0,1}, qty U[1,50], price = qty*U[900,105000]/100,
    disc in {0.00..0.10}, tax in {0.00..0.08}, d U[0,2556)."""
    r = randint(seed, 1, n, 0, 4, np.int64)
    f1 = np.where(r >= 2, 2, r).astype(np.int8)
    f2 = randint(seed, 2, n, 0, 2, np.int8)
    qty = randint(seed, 3, n, 1, 51, np.int64)
    price = qty.astype(np.float64) * randint(seed, 4, n, 900, 105001, np.int64).astype(np.float64) / 100.0
    disc = randint(seed, 5, n, 0, 11, np.int64).astype(np.float64) / 100.0
    tax = randint(seed, 6, n, 0, 9, np.int64).astype(np.float64) / 100.0
    d = randint(seed, 7, n, 0, 2556, np.int32)
    return {"rowid": np.arange(n, dtype=np.int64), "f1": f1, "f2": f2, "qty": qty,
            "price": price, "disc": disc, "tax": tax, "d": d}


def permutation(seed: int, col: int, n: int) -> np.ndarray:
    """Seeded permutation of [0, n): rows sorted by their counter hash."""
    return np.argsort(u64(seed, col, n), kind="stable").astype(np.int64)


def join_tables(n_build: int, n_probe: int, seed: int = 11):
    """Hash-join workload: build bk = permutation of [0,n_build), bd, bf;
    probe pk uniform over build keys, pv U[0,1), pq U[1,50]."""
    build = {"bk": permutation(seed, 1, n_build),
             "bd": randint(seed, 2, n_build, 0, 2556, np.int32),
             "bf": randint(seed, 3, n_build, 0, 5, np.int8)}
    probe = {"rowid": np.arange(n_probe, dtype=np.int64),
             "pk": randint(seed, 4, n_probe, 0, n_build, np.int64),
             "pv": uniform(seed, 5, n_probe),
             "pq": randint(seed, 6, n_probe, 1, 51, np.int64)}
    return build, probe


def lognormal_sizes(n: int, seed: int, median: int = 262_144, sigma: float = 1.2,
                    lo: int = 4096, hi: int = 8_388_608) -> np.ndarray:
    """Seeded log-normal row counts clipped to [lo, hi] (Box-Muller over two
    counter-based uniforms). With 64 batches and seed 5: 43.8M rows, median
    266K, from 15K to 7.9M."""
    u1 = uniform(seed, 1, n)
    u2 = uniform(seed, 2, n)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    return np.clip(np.rint(median * np.exp(sigma * z)), lo, hi).astype(np.int64)


def poisson_arrivals(n: int, qps: float, seed: int = 23) -> np.ndarray:
    """Arrival times in microseconds of a seeded Poisson process."""
    gaps = -np.log1p(-uniform(seed, 1, n)) / qps * 1e6
    return np.cumsum(gaps) - gaps[0]


def dim_table(seed: int = 31, days: int = 2556) -> dict:
    """Date dimension: dk = 0..days-1, dflag uniform in [0, 5)."""
    return {"dk": np.arange(days, dtype=np.int64), "dflag": randint(seed, 1, days, 0, 5, np.int8)}
