"""reffine's PageRank step (benchmark/include/misc.h:212-291) as written:
a named Op with an index built over it (`_buildidx`) is computed first and
read as a table by the fold of every node's in-edges. Against
oracle.graph on the same seeded graph: node keys exact, float32 ranks bit
for bit (every run folded in row order, one float32 operation at a time)."""
import numpy as np
import pytest

from oracle import graph as og
from reffine_amd import DeviceTable, compile_op, datagen, graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m", [(1000, 9000), (20000, 300000)])
def test_pagerank_step(device, n, m):
    t = datagen.graph_tables(n, m, seed=5)
    rb = graph.record_batches(t)
    E = DeviceTable.from_arrow(rb["edges"], dim=2).build_index()
    R = DeviceTable.from_arrow(rb["rev_edges"], dim=2).build_index()
    P = DeviceTable.from_arrow(rb["pr"]).build_index()
    q = compile_op(graph.pagerank(n))
    out = q(E, R, P)
    k, v = og.pagerank(t, n)
    assert out.nrows == len(k)
    np.testing.assert_array_equal(out.column(0), k)
    np.testing.assert_array_equal(out.column(1).view(np.uint32), v.view(np.uint32))
    again = q(E, R, P)
    np.testing.assert_array_equal(again.column(1).view(np.uint32), v.view(np.uint32))
    assert q.time(E, R, P, iters=2) > 0


@pytest.mark.parametrize("n", [64, 256])
def test_nbody_step(device, n):
    """reffine's Nbody step (misc.h:63-210) as written: the pair table (an
    Op over two ranges reading bodies at computed rows) and the per-body
    window sums over it are computed first, the last Op reads them at its
    iterator. Float64 results bit for bit against oracle.graph.nbody."""
    b = datagen.bodies(n, seed=3)
    T = DeviceTable.from_columns([b[c] for c in ("id", "m", "x", "y", "z", "vx", "vy", "vz")],
                                 dim=1)
    out = compile_op(graph.nbody(n))(T)
    want = og.nbody(b, n)
    assert out.nrows == n
    for c in range(8):
        got = out.column(c)
        np.testing.assert_array_equal(got.view(np.uint64), np.asarray(want[c]).view(np.uint64),
                                      err_msg=f"column {c}")
