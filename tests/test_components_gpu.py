"""Functions whose components cannot share one scan -- reductions over
different tables, a sequential fold beside another component -- run as one
query per component on the same inputs (the reference emits each component
as a loop of its own, src/pass/loopgen.cpp:113-269), the output evaluated
over their results; checked against row loops and the CPU JIT restatement."""
import math

import numpy as np
import pytest

import seq_queries as sq
from oracle import cjit
from reffine_amd import DeviceTable, compile_op

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(-1000, 1000, n).astype(np.int64), rng.random(n)


def test_reductions_over_two_tables(device):
    k, v = _data(300_000, 1)
    x = np.random.default_rng(2).random(123_457)
    T = DeviceTable.from_columns([np.arange(len(v)), k, v], dim=2)
    U = DeviceTable.from_columns([np.arange(len(x)), x], dim=2)
    q = compile_op(sq.sum_and_count_of_two_tables())
    assert "independent components" in q.describe()
    s, c = q(T, U)
    assert c == len(x)
    assert abs(s - math.fsum(v)) <= 1e-12 * abs(math.fsum(v))


@pytest.mark.parametrize("n", [1, 64, 100_003])
def test_sequential_fold_beside_a_sum(device, n):
    k, v = _data(n, n)
    T = DeviceTable.from_columns([np.arange(n), k, v], dim=2)
    d, s = compile_op(sq.decay_and_sum())(T)
    (want_d,) = sq.loop("decay", k, v)
    assert np.float64(d).view(np.uint64) == np.float64(want_d).view(np.uint64)
    (cd, cs) = cjit.compile_op(sq.decay_and_sum())({"cols": [np.arange(n), k, v]})
    assert np.float64(d).view(np.uint64) == np.float64(cd[0]).view(np.uint64)
    assert abs(s - math.fsum(v)) <= 1e-12 * abs(math.fsum(v))


def test_selection_beside_a_sequential_fold(device):
    n = 200_000
    k, v = _data(n, 5)
    T = DeviceTable.from_columns([np.arange(n, dtype=np.int64), k, v])
    q = compile_op(sq.selection_and_decay())
    for _ in range(2):  # a second run reuses the parts' workspaces
        sel, d = q(T)
        m = v > 0.9
        assert np.array_equal(sel.column(0), np.arange(n)[m])
        assert np.array_equal(sel.column(1), v[m])
        s = 0.25
        for x in v:
            s = s * 0.5 + float(x)
        assert np.float64(d).view(np.uint64) == np.float64(s).view(np.uint64)
