"""Autotuner for lowered queries: launch shape and hash-index load factor.

Sweeps launch shapes (BLOCK, ITEMS, SUB, WAVES_PER_CU, PIPELINE, LSUB) for a
query on its real input tables, times each candidate's kernels with HIP
events (Query.time) and keeps the fastest; for a group join it then sweeps
the load factor of the hash index over its build keys (index_scale: slots
per key) at the chosen shape. Choices persist per query shape -- the IR with
symbol ids and constant values removed, the mode, and each input's row count
rounded to a power of two, so one tuning serves every batch of a similar
size -- in a JSON file ($RF_TUNING_CACHE, default
~/.cache/reffine_amd/tuning.json). `tuned()` compiles from that file, tuning
first on a miss; the batched and served paths run whatever Query they are
given, so queries compiled by `tuned()` run there with their tuned shapes.
Deterministic queries are tuned like the others; their reduction order then
follows the chosen shape (Query.tuning() reports it).
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import re
import threading

from .runtime import Query, ReffineError, compile_op

_lock = threading.Lock()
_mem: dict[str, dict] = {}


def cache_path() -> str:
    return os.environ.get("RF_TUNING_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "reffine_amd", "tuning.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict) -> None:
    with _lock:
        _mem[key] = entry
        d = _load()
        d[key] = entry
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def lookup(key: str) -> dict | None:
    """The persisted choice for a shape key (memory first, then the file)."""
    if key in _mem:
        return _mem[key]
    e = _load().get(key)
    if e is not None:
        _mem[key] = e
    return e


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


INDEX_SCALES = (2, 4, 8)


def shape_key(q: Query, inputs) -> str:
    text = re.sub(r"#\d+", "", q.text)  # symbol ids
    text = re.sub(r"\(const (\w+) [^)]*\)", r"(const \1)", text)  # constant values
    rows = []
    for x in inputs:
        n = getattr(x, "nrows", None)
        rows.append(-1 if n is None else (0 if n <= 0 else int(round(math.log2(n)))))
    h = hashlib.sha1(text.encode()).hexdigest()[:20]
    return f"{h}:{int(q.deterministic)}:{','.join(map(str, rows))}"


def _is_group_join(q: Query) -> bool:
    return "group join:" in q.describe()


def autotune(func, inputs, deterministic: bool = False, iters: int = 10,
             cands: list[tuple[int, ...]] | None = None, scales=None, verbose: bool = False,
             persist: bool = True):
    """Returns (best Query, {(tuning, index_scale): ms}) for `func` on
    `inputs`; the choice is persisted under the query's shape key."""
    base = compile_op(func, deterministic=deterministic)
    key = shape_key(base, inputs)
    has_sel = "(op " in base.text and "\n  select " in base.describe()
    gj = _is_group_join(base)
    results = {}
    best, best_ms, best_t = None, float("inf"), None
    for tn in cands or candidates(has_sel):
        try:
            q = compile_op(func, deterministic=deterministic, tuning=tn)
        except ReffineError:
            continue
        q(*inputs)
        ms = q.time(*inputs, iters=iters)
        results[(tn, None)] = ms
        if verbose:
            print(f"  tuning {tn}: {ms:.4f} ms", flush=True)
        if ms < best_ms:
            best, best_ms, best_t = q, ms, tn
    if best is None:
        return base, results
    best_scale = None
    if gj:
        for sc in (scales or INDEX_SCALES):
            q = compile_op(func, deterministic=deterministic, tuning=best_t, index_scale=sc)
            q(*inputs)
            ms = q.time(*inputs, iters=iters)
            results[(best_t, sc)] = ms
            if verbose:
                print(f"  tuning {best_t} index scale {sc}: {ms:.4f} ms", flush=True)
            if ms < best_ms:
                best, best_ms, best_scale = q, ms, sc
    if persist:
        _store(key, {"tuning": list(best.tuning()), "index_scale": best_scale,
                     "ms": round(best_ms, 5), "default_ms": None})
    return best, results


def tuned(func, inputs, deterministic: bool = False, iters: int = 10,
          verbose: bool = False) -> Query:
    """The query compiled with its persisted tuning for inputs of this shape,
    tuning first (and persisting) on a miss."""
    base = compile_op(func, deterministic=deterministic)
    e = lookup(shape_key(base, inputs))
    if e is None:
        q, _ = autotune(func, inputs, deterministic=deterministic, iters=iters, verbose=verbose)
        return q
    return compile_op(func, deterministic=deterministic, tuning=tuple(e["tuning"]),
                      index_scale=e.get("index_scale"))
