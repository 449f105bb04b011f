"""Batched and served execution against the plain per-query path.

Each unit of a batch runs the deterministic scan's per-tile fold and fixed
combine order, so its output is bit-identical to the same deterministic
query run alone; selections are identical row for row.
"""
from __future__ import annotations

import numpy as np
import pytest

from reffine_amd import DeviceTable, compile_op, datagen, workloads
from reffine_amd.batch import Batch, reduction_bytes, serve, serve_plain

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 100, 4096, 70_001, 300_000, 1_200_000]


@pytest.fixture(scope="module")
def batches(device):
    out = []
    for i, n in enumerate(SIZES):
        c = datagen.group_table(n, 100 + i)
        t1 = DeviceTable.from_columns([c["rowid"], c["f1"], c["f2"], c["qty"], c["price"],
                                       c["disc"], c["tax"], c["d"]], dim=1)
        out.append((t1, t1.view(2)))
    return out


@pytest.fixture(scope="module")
def queries(device):
    return {"q6": compile_op(workloads.q6_template(), deterministic=True),
            "q1": compile_op(workloads.group_agg_c3(None), deterministic=True),
            "join": compile_op(workloads.dim_join_template(), deterministic=True)}


@pytest.fixture(scope="module")
def dim(device):
    d = datagen.dim_table()
    return d, DeviceTable.from_columns([d["dk"], d["dflag"]], dim=1)


BATCHED = ("q6", "q1")  # templates a batch runner takes


def _work(queries, batches, mix, D=None):
    work = []
    for name, b, params in mix:
        t1, t2 = batches[b]
        if name == "join":
            work.append((queries[name], (D, t1) + params))
        else:
            work.append((queries[name], (t1 if name == "q6" else t2,) + params))
    return work


def _plain(work):
    return [q(*inputs) for q, inputs in work]


def _cols(t):
    """A selection as host columns (plain results hold device tables)."""
    return [t.column(c) for c in range(t.ncols)] if hasattr(t, "column") else t


def _same(a, b):
    """Outputs equal bit for bit (floats compared by their bits)."""
    a, b = _cols(a), _cols(b)
    if isinstance(a, tuple):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, list):  # selection columns
        return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))
    if isinstance(a, float):
        return np.float64(a).view(np.int64) == np.float64(b).view(np.int64)
    return a == b


def _cmp(plain, batched):
    assert len(plain) == len(batched)
    for i, (p, b) in enumerate(zip(plain, batched)):
        assert _same(p, b), f"unit {i} differs"


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_plain(queries, batches, graph):
    mix = workloads.c5_mix(60, len(SIZES), seed=3, templates=BATCHED)
    work = _work(queries, batches, mix)
    plain = _plain(work)
    b = Batch(max_units=64, arena_bytes=1 << 30, use_graph=graph)
    _cmp(plain, b.run(work))
    # replays of the captured round, with eager work in between
    for rep in range(3):
        _cmp(plain, b.run(work))
        _cmp(plain, _plain(work))


def test_batch_admission_by_footprint(queries, batches):
    """A small arena admits a few units per round; every unit still runs and
    matches the plain result."""
    mix = workloads.c5_mix(24, len(SIZES), seed=4, templates=BATCHED)
    work = _work(queries, batches, mix)
    big = max(Batch(4, 1 << 20).footprint(q, *inp) for q, inp in work)
    b = Batch(max_units=64, arena_bytes=big + (big // 2), use_graph=False)
    assert b.add(*_flat(work[0])) is not None
    b.reset()
    _cmp(_plain(work), b.run(work))


def _flat(w):
    return (w[0],) + tuple(w[1])


def test_batch_same_query_many_units(queries, batches):
    """Units of one kernel over every batch size in one launch."""
    work = [(queries["q6"], (t1, 100 * i, 0.01, 0.05, 30)) for i, (t1, _) in enumerate(batches)]
    work += [(queries["q1"], (t2, 2000 + i)) for i, (_, t2) in enumerate(batches)]
    _cmp(_plain(work), Batch(max_units=32, use_graph=True).run(work))


@pytest.mark.parametrize("runners", [0, 2, 4])
def test_serve_matches_plain_and_reports_latency(queries, batches, dim, runners):
    """Served results (2, 3 -- the default -- or 4 rounds in flight) equal
    the plain runs of the same requests."""
    mix = workloads.c5_mix(400, len(SIZES), seed=5)
    assert {m[0] for m in mix} == {"q6", "q1", "join"}
    work = _work(queries, batches, mix, dim[1])
    arrivals = datagen.poisson_arrivals(len(work), qps=20_000.0, seed=6)
    reqs = [(q, inp, t) for (q, inp), t in zip(work, arrivals)]
    caps = [int(inp[0].nrows) for _, inp, _ in reqs]
    s = serve(reqs, max_units=128, arena_bytes=1 << 30, use_graph=True, sel_cap_rows=caps,
              runners=runners)
    p = serve_plain(reqs, sel_cap_rows=caps)
    assert (s.status == 0).all() and (p.status == 0).all()
    assert (s.latency_us >= 0).all() and (p.latency_us >= 0).all()
    assert np.array_equal(s.out_rows, p.out_rows)
    for i, (q, _, _) in enumerate(reqs):
        assert reduction_bytes(q, s.outputs[i]) == reduction_bytes(q, p.outputs[i]), i
        assert s.selections[i] == p.selections[i], i


def test_serve_rejects_what_cannot_run(queries, batches):
    """A unit larger than the arena fails alone; the others complete."""
    t1, t2 = batches[-1]
    small_t1, _ = batches[3]
    reqs = [(queries["q6"], (small_t1, 0, 0.0, 0.1, 50), 0.0),
            (queries["q6"], (t1, 0, 0.0, 0.1, 50), 1.0),
            (queries["q1"], (batches[3][1], 2500), 2.0)]
    s = serve(reqs, max_units=8, arena_bytes=1 << 20, use_graph=False)
    assert s.status[0] == 0 and s.status[2] == 0
    assert s.status[1] != 0


@pytest.mark.parametrize("graph", [False, True])
def test_batch_group_joins_match_plain(queries, batches, dim, graph):
    """Group join units (several build flags, a flag no day has, empty and
    large probe batches) in rounds with the other templates: each unit's
    table equals its plain run bit for bit."""
    mix = workloads.c5_mix(48, len(SIZES), seed=8)
    work = _work(queries, batches, mix, dim[1])
    work += [(queries["join"], (dim[1], batches[b][0], f, 30)) for b, f in
             [(0, 1), (6, 2), (5, 7), (2, 0), (6, 7)]]
    plain = _plain(work)
    b = Batch(max_units=24, arena_bytes=1 << 30, use_graph=graph)
    for rep in range(2):
        _cmp(plain, b.run(work))


def test_dim_join_against_oracle_and_build_cache(queries, batches, dim):
    """The join template over several batches and flags (the build side is
    kept while the build table and flag repeat) against the oracle."""
    from oracle import brute
    d, D = dim
    q = queries["join"]
    for i, f in ((5, 1), (6, 1), (6, 3), (5, 3), (3, 3)):
        n = SIZES[i]
        c = datagen.group_table(n, 100 + i)
        out = q(D, batches[i][0], f, 30)
        keys, sums, counts = brute.dim_join(d, c, f, 30)
        assert np.array_equal(out.column(0), keys)
        assert np.array_equal(out.column(2), counts)
        assert np.array_equal(out.column(1).view(np.int64), sums.view(np.int64))
