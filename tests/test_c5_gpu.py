"""The batched/served configuration at full size: 64 device-resident record
batches of log-normal row counts (43.8M rows in all), the 4,096-query mix
drawn uniformly from three templates (a Q6-shaped scan-filter-aggregate with
its selection, a Q1-shaped 6-group aggregate, a hash join with the date
dimension grouped by day), with seeded constants.

* Served at a fixed Poisson rate through rf_serve (continuous batching over
  two runners), every request's result equals its plain run (rf_query_run,
  the queries compiled in the deterministic mode) bit for bit: reduction
  fields, selection rows, joined keys.
* A seeded sample of 64 units per template (batches of at most 2M rows, where
  the oracle's correctly rounded sums stay quick) matches the oracle: counts,
  integer sums, selections and group keys exactly, FP64 sums within 1e-12 of
  math.fsum, and the join's per-day sums bit for bit (its fold runs in row
  order in the deterministic mode).
"""
import struct

import numpy as np
import pytest

import bench
from oracle import brute
from reffine_amd import batch, datagen, workloads

pytestmark = pytest.mark.gpu

N_QUERIES = 4096
RATE_QPS = 40_000.0


def test_c5_served_equals_plain_and_oracle(device):
    sizes, tables, queries, work, _ = bench.c5_setup(N_QUERIES)
    mix = workloads.c5_mix(N_QUERIES, len(sizes), seed=17)  # the mix c5_setup drew
    assert len(mix) == len(work)
    caps = [int(0.05 * sizes[b]) + 256 if name == "q6" else (2600 if name == "join" else 0)
            for name, b, _ in mix]
    arrivals = datagen.poisson_arrivals(len(work), RATE_QPS, seed=29)
    served = batch.serve([(q, inp, t) for (q, inp), t in zip(work, arrivals)], sel_cap_rows=caps)
    plain = batch.serve_plain([(q, inp, 0.0) for q, inp in work], sel_cap_rows=caps)
    assert (served.status == 0).all() and (plain.status == 0).all()
    assert np.array_equal(served.out_rows, plain.out_rows)
    for i, (q, _) in enumerate(work):
        assert batch.reduction_bytes(q, served.outputs[i]) == \
            batch.reduction_bytes(q, plain.outputs[i]), i
        assert served.selections[i] == plain.selections[i], i
    lat = served.latency_us
    assert (lat > 0).all() and np.isfinite(np.percentile(lat, 99))

    # Oracle sample: 64 units per template over batches of at most 2M rows.
    rng = np.random.default_rng(23)
    host = {}
    dim = datagen.dim_table()
    checked = {"q6": 0, "q1": 0, "join": 0}
    for name in checked:
        idx = [i for i, (nm, b, _) in enumerate(mix) if nm == name and sizes[b] <= 2_000_000]
        for i in rng.choice(idx, size=64, replace=False):
            _, b, params = mix[i]
            if b not in host:
                host[b] = datagen.group_table(int(sizes[b]), 1000 + b)
            c = host[b]
            q, inp = work[i]
            if name == "q6":
                cnt, sq, spd, rows = brute.q6_batch(c, *params)
                got = struct.unpack("<qqd", batch.reduction_bytes(q, plain.outputs[i]))
                assert got[0] == cnt and got[1] == sq
                assert brute.rel_err(got[2], spd) <= 1e-12
                assert plain.out_rows[i] == cnt
                assert np.frombuffer(plain.selections[i], np.int64).tolist() == rows.tolist()
            elif name == "q1":
                assert bench._c3_ok(q.decode(plain.outputs[i]), brute.group_c3(c, params[0]))
            else:
                out = q(*inp)
                keys, sums, counts = brute.dim_join(dim, c, *params)
                assert np.array_equal(out.column(0), keys)
                assert np.array_equal(out.column(2), counts)
                assert np.array_equal(out.column(1).view(np.int64), sums.view(np.int64))
            checked[name] += 1
    assert checked == {"q6": 64, "q1": 64, "join": 64}
