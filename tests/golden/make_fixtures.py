"""Writes the golden fixtures the reference's own tests use (run in a
checkout that has the reference tree; the output is committed).

  students.npz   the int64 columns of benchmark/lib/students.csv that
                 benchmark/lib/students.py:5-12,41-44 writes to
                 students.arrow (the file test/src/test_reduce.cpp and
                 test/src/test_transform.cpp read): id, hours_studied,
                 attendance, sleep_hours, previous_scores, exam_score
  runend.npz     the arrays of benchmark/lib/runend.py:5-12 (x is run-end
                 encoded in runend.arrow; stored here decoded)
  expected.json  values the reference tests assert
"""
import csv
import json
import os
import sys

import numpy as np

REF = os.environ.get("REFFINE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
COLS = ["id", "hours_studied", "attendance", "sleep_hours", "previous_scores", "exam_score"]


def main() -> int:
    path = os.path.join(REF, "benchmark", "lib", "students.csv")
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    data = {c: np.array([int(r[c]) for r in rows], dtype=np.int64) for c in COLS}
    np.savez_compressed(os.path.join(HERE, "students.npz"), **data)
    np.savez_compressed(os.path.join(HERE, "runend.npz"),
                        x=np.array([1, 1, 1, 2, 2, 3, 3, 3, 3, 4], dtype=np.int64),
                        y=np.array([1, 3, 7, 1, 9, 2, 3, 7, 8, 3], dtype=np.int64),
                        d=np.arange(10, dtype=np.int64))
    expected = {
        # test/src/test_reduce.cpp:58 -- sum of hours_studied over all rows
        "aggregate_loop_test": 131977,
        # test/src/test_reduce.cpp:121 -- sum of hours_studied, 10 <= id <= 48
        "aggregate_op_test": 696,
        # test/src/test_transform.cpp:43 -- rows with id > 5
        "transform_op_test_len": len(rows) - 5,
        # test/src/test_join.cpp:62-86 -- [0,10) & [5,15) & [7,12)
        "join_op_test_keys": [7, 8, 9],
        # test/src/test_multidim.cpp:42-54 -- per-run sums of d
        "multidim_op_test": {"keys": [1, 2, 3, 4], "sums": [3, 7, 26, 9]},
        "students_rows": len(rows),
    }
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)
    print(f"wrote fixtures for {len(rows)} student rows")
    return 0


if __name__ == "__main__":
    sys.exit(main())
