"""The C ABI from a C host program (no Python in the process): upload
through the Arrow C data interface, compile from IR text, run, compare with
a C loop. The program is tests/c_abi/rf_c_scan.c, built by
reffine_amd.build.build_c_abi_test (run by __graft_entry__.build)."""
from __future__ import annotations

import os
import subprocess

import pytest

from reffine_amd import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "c_abi", "_build", "rf_c_scan")


@pytest.mark.parametrize("lo,hi,vmax,rows", [(365, 730, 0.5, 1 << 20), (0, 2556, 1.0, 1000),
                                             (100, 100, 0.5, 5000)])
def test_c_host_program(tmp_path, lo, hi, vmax, rows):
    assert os.path.exists(EXE), "tests/c_abi/_build/rf_c_scan is missing: run the build first"
    ir = tmp_path / "c1.ir"
    ir.write_text(workloads.scan_agg_c1(lo, hi, vmax).ir())
    r = subprocess.run([EXE, str(ir), str(lo), str(hi), str(vmax), str(rows)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "c abi: rows" in r.stdout
