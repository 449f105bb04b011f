"""Launch-shape autotuner for lowered queries.

Sweeps launch shapes (BLOCK, ITEMS, SUB, WAVES_PER_CU, PIPELINE, LSUB) for a
query on its real input tables, times each
candidate's kernel with HIP events (Query.time) and keeps the fastest. The
choice is cached in memory per (function IR shape, table row counts).
Deterministic queries are tuned like the others; their reduction order then
follows the chosen shape (Query.tuning() reports it).
"""
from __future__ import annotations

import re

from .runtime import Query, ReffineError, compile_op

_cache: dict[tuple, tuple[int, ...]] = {}


def candidates(has_selection: bool) -> list[tuple[int, int, int, int, int, int]]:
    """Launch shapes to try (BLOCK, ITEMS, SUB, WAVES_PER_CU, PIPELINE, LSUB):
    wave-tiles of SUB*64*ITEMS rows, up to 32 rows per lane per tile, loaded
    LSUB sub-steps at a time (0: as many as the register budget allows)."""
    out = []
    for items in (2, 4):
        for sub in (2, 4, 8, 16):
            if items * sub > 32:
                continue
            for wpc in (8, 16):
                for pipe in (0, 1):
                    for lsub in ((0, 4) if sub > 4 else (0,)):
                        out.append((256, items, sub, wpc, pipe, lsub))
    return out


def _shape_key(q: Query, inputs) -> tuple:
    text = re.sub(r"#\d+", "", q.text)  # symbol ids
    text = re.sub(r"\(const (\w+) [^)]*\)", r"(const \1)", text)  # constant values
    rows = tuple(getattr(x, "nrows", x) for x in inputs)
    return (text, q.deterministic, rows)


def autotune(func, inputs, deterministic: bool = False, iters: int = 10,
             cands: list[tuple[int, ...]] | None = None, verbose: bool = False):
    """Returns (best Query, {tuning: ms}) for `func` on `inputs`."""
    base = compile_op(func, deterministic=deterministic)
    key = _shape_key(base, inputs)
    if key in _cache and cands is None:
        return compile_op(func, deterministic=deterministic, tuning=_cache[key]), {}
    has_sel = "(op " in base.text and _output_has_op(base)
    results = {}
    best, best_ms = None, float("inf")
    for tn in cands or candidates(has_sel):
        try:
            q = compile_op(func, deterministic=deterministic, tuning=tn)
        except ReffineError:
            continue
        q(*inputs)
        ms = q.time(*inputs, iters=iters)
        results[tn] = ms
        if verbose:
            print(f"  tuning {tn}: {ms:.4f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = q, ms
    if best is None:
        return base, results
    _cache[key] = best.tuning()
    return best, results


def _output_has_op(q: Query) -> bool:
    d = q.describe()
    return "\n  select " in d
