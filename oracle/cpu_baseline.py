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
