"""CPU baseline timing of the oracle's restatement (test infrastructure,
used only by bench.py's cpu_baseline leg)."""
from __future__ import annotations

import os
import time

import numpy as np


def _host() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_c2(rows: int, seed: int = 2, reps: int = 3) -> dict:
    """Times the scan-select query on `rows` rows of the same table layout
    with the CPU JIT restatement when built, else the numpy restatement."""
    from reffine_amd import datagen
    from . import brute
    c = datagen.scan_table(rows, seed, with_q=True, rowid_dtype=np.int32)
    try:
        from . import cjit
        fn = cjit.c2_query()
        kind, cores = "port", 1
        run = lambda: fn(c["rowid"], c["v"], c["d"], c["q"])  # noqa: E731
        what = "oracle/cjit: CEmitter-style C, clang -O3, single thread"
    except Exception:  # cjit not built on this host
        run = lambda: brute.scan_c2(c["v"], c["d"], c["q"], rowid=c["rowid"])  # noqa: E731
        kind, cores = "port", 1
        what = "oracle/brute: numpy restatement, single thread"
    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    s = float(np.median(times))
    gbs = 20.0 * rows / s / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": cores, "kind": kind,
            "sample": f"{rows} rows of the same 5-column layout, median of {reps} runs; {what}; "
                      f"host {_host()} ({os.cpu_count()} logical CPUs)",
            "rows_per_s": round(rows / s)}


def time_c1(rows: int = 1_048_576, seed: int = 42, reps: int = 7) -> dict:
    """The C1 configuration itself: COUNT(*), SUM(v*k) WHERE 365 <= d < 730
    AND v < 0.5 over 1,048,576 rows, through the CPU JIT restatement
    (CEmitter-style C, clang -O3) on one thread; median of `reps` runs
    after one warm-up, the query call alone timed (compile excluded)."""
    from reffine_amd import datagen, workloads
    from . import cjit
    c = datagen.scan_table(rows, seed, with_q=False, rowid_dtype=np.int64)
    fn = cjit.compile_op(workloads.scan_agg_c1())
    inp = {"cols": [c["rowid"], c["k"], c["v"], c["d"]]}
    s = _median_time(lambda: fn(inp), reps)
    return {"value": round(20.0 * rows / s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "ms_per_query": round(s * 1e3, 4), "rows_per_s": round(rows / s),
            "sample": f"the whole C1 workload: {rows} rows (rowid, k i64, v f64, d i32), median of "
                      f"{reps} runs; oracle/cjit: CEmitter-style C, clang -O3, single thread; "
                      f"host {_host()} ({os.cpu_count()} logical CPUs)"}


def _median_time(run, reps: int) -> float:
    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def time_c3(rows: int, seed: int = 7, reps: int = 3) -> dict:
    """Times the grouped aggregate (TPC-H Q1 idiom) with the CPU JIT
    restatement on `rows` rows of the same 8-column layout."""
    from reffine_amd import datagen, workloads
    from . import cjit
    c = datagen.group_table(rows, seed)
    fn = cjit.compile_op(workloads.group_agg_c3())
    inp = {"cols": [c["rowid"], c["f1"], c["f2"], c["qty"], c["price"], c["disc"], c["tax"],
                    c["d"]]}
    s = _median_time(lambda: fn(inp), reps)
    return {"value": round(38.0 * rows / s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{rows} rows of the same 8-column layout, median of {reps} runs; oracle/cjit: "
                      f"CEmitter-style C, clang -O3, single thread; host {_host()} "
                      f"({os.cpu_count()} logical CPUs)",
            "rows_per_s": round(rows / s)}


def time_c4(n_build: int, n_probe: int, seed: int = 11, reps: int = 3) -> dict:
    """Times the hash join + grouped aggregate with the numpy restatement
    (sort the build keys, binary-search each probe key, sequential per-group
    fold) on a sample; the reference's own CPU form is a nested loop over
    build x probe rows, far slower at these sizes."""
    from reffine_amd import datagen
    from . import brute
    b, p = datagen.join_tables(n_build, n_probe, seed)
    s = _median_time(lambda: brute.group_join_c4(b, p), reps)
    keys = int((b["bf"] == 1).sum())
    nbytes = 9 * n_build + 24 * n_probe + 24 * keys
    return {"value": round(nbytes / s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"build {n_build} / probe {n_probe} rows of the same layout, median of {reps} "
                      f"runs; oracle/brute: numpy sort + searchsorted + sequential add.at, single "
                      f"thread; host {_host()} ({os.cpu_count()} logical CPUs)",
            "rows_per_s": round(n_probe / s)}
