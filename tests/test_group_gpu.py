"""GPU parity of the low-cardinality grouped aggregate (TPC-H Q1 idiom,
6 groups x 6 fields) against the brute-force oracle; deterministic mode
bit-exact against the fixed-order restatement."""
import numpy as np
import pytest

from oracle import brute, cjit, detorder
from reffine_amd import DeviceTable, compile_op, datagen, workloads

pytestmark = pytest.mark.gpu
COLS = ["rowid", "f1", "f2", "qty", "price", "disc", "tax", "d"]


def c3_table(n, seed=7):
    c = datagen.group_table(n, seed)
    return c, DeviceTable.from_columns([c[k] for k in COLS], dim=2)


def check(got, want):
    for i, (g, w) in enumerate(zip(got, want)):
        if isinstance(w, int):
            assert g == w, (i, g, w)
        else:
            assert brute.rel_err(g, w) <= 1e-12, (i, g, w)


@pytest.mark.parametrize("n", [0, 10, 4099, 1 << 20])
def test_c3_group_agg(device, n):
    c, t = c3_table(n)
    got = compile_op(workloads.group_agg_c3())(t)
    check(got, brute.group_c3(c))


def test_c3_deterministic_bit_exact(device):
    n = 500_000
    c, t = c3_table(n, seed=8)
    q = compile_op(workloads.group_agg_c3(), deterministic=True)
    got = q(t)
    block, items, sub = q.tuning()[:3]
    keep = c["d"] <= 2526
    vals = brute.group_c3_contribs(c)
    i = 0
    for a, b in workloads.C3_GROUPS:
        m = keep & (c["f1"] == a) & (c["f2"] == b)
        for x in vals + [np.ones(n, np.int64)]:
            want = detorder.det_sum(x, m, 0, n, block, items, sub)
            assert got[i] == want.item(), (i, got[i], want)
            i += 1
    assert q(t) == got


def test_c3_deterministic_equals_cjit_fixed_order(device):
    """The deterministic mode's 36 Q1 fields equal the CPU JIT
    restatement's fixed-order emission at the kernel's launch shape, bit
    for bit."""
    n = 700_001
    c, t = c3_table(n, seed=10)
    q = compile_op(workloads.group_agg_c3(), deterministic=True)
    (want,) = cjit.compile_op(workloads.group_agg_c3(), fixed_order=q.tuning()[:3])(
        {"cols": [c[k] for k in COLS]})
    got = q(t)
    assert [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in got] == \
        [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in want]


def test_c3_full_size_60m(device):
    n = 60_000_000
    c, t = c3_table(n, seed=9)
    got = compile_op(workloads.group_agg_c3())(t)
    check(got, brute.group_c3(c))
    assert sum(got[5::6]) == int((c["d"] <= 2526).sum())
