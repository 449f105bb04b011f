"""Union merges on the GPU (an Op over _in(t, A) | _in(t, B), the
reference's UnionSpace, src/iter/iter_space.cpp:216-316, as lowered by
src/pass/loopgen.cpp:44-110): the outer join of
benchmark/include/micro.h:79-95 against the oracle's merge (oracle/merge.py),
keys and int64 outputs bit for bit."""
import numpy as np
import pytest

import merge_queries as mq
from oracle import merge
from reffine_amd import DeviceTable, compile_op
from reffine_amd.runtime import ReffineError

pytestmark = pytest.mark.gpu


def _keys(rng, n, span):
    if n == 0:
        return np.zeros(0, np.int64)
    return np.sort(rng.choice(span, size=n, replace=False)).astype(np.int64) - span // 3


def _check(q, ka, va, kb, vb, valid_a=None, valid_b=None):
    ta = DeviceTable.from_columns([ka, va], validity=[valid_a, None])
    tb = DeviceTable.from_columns([kb, vb], validity=[valid_b, None])
    out = q(ta, tb)
    t, d = merge.ojoin(ka, va, kb, vb, valid_a, valid_b)
    assert out.nrows == len(t)
    assert np.array_equal(out.column(0), t)
    assert np.array_equal(out.column(1), d)


@pytest.mark.parametrize("na,nb", [(0, 0), (0, 5), (5, 0), (1, 1), (7, 9), (100, 37),
                                   (4097, 1), (100_000, 80_000)])
def test_outer_join_against_oracle(device, na, nb):
    rng = np.random.default_rng(na * 7 + nb)
    span = max(2 * (na + nb), 4)
    ka, kb = _keys(rng, na, span), _keys(rng, nb, span)
    va = rng.integers(-10**12, 10**12, na)
    vb = rng.integers(-10**12, 10**12, nb)
    _check(compile_op(mq.ojoin()), ka, va, kb, vb)


def test_outer_join_same_table_on_both_sides(device):
    """micro.h joins fake_data with itself: every key once, diff 0."""
    rng = np.random.default_rng(3)
    k = _keys(rng, 300_000, 1_000_000)
    v = rng.integers(-1000, 1000, len(k))
    t = DeviceTable.from_columns([k, v])
    out = compile_op(mq.ojoin())(t, t)
    assert np.array_equal(out.column(0), k)
    assert not out.column(1).any()


def test_outer_join_disjoint_and_exhausted_sides(device):
    """Keys of A all below B's (A runs out first: its rows read 0), and
    interleaved with long runs of one side."""
    q = compile_op(mq.ojoin())
    ka = np.arange(0, 5000, dtype=np.int64)
    kb = np.arange(10_000, 12_000, dtype=np.int64)
    va = np.arange(5000, dtype=np.int64) + 1
    vb = -np.arange(2000, dtype=np.int64) - 1
    _check(q, ka, va, kb, vb)
    _check(q, kb, vb, ka, va)
    ka = np.concatenate([np.arange(0, 3000), np.arange(9000, 9100)]).astype(np.int64)
    kb = np.arange(2990, 9050, 2, dtype=np.int64)
    _check(q, ka, np.ones(len(ka), np.int64), kb, np.full(len(kb), 5, np.int64))


def test_outer_join_null_keys_and_filter(device):
    """Null key rows still take part in the merge but emit only when the
    other side's current row is a valid key (UnionSpace::_iter_cond); the
    filtered form keeps even keys only."""
    rng = np.random.default_rng(5)
    ka, kb = _keys(rng, 50_000, 200_000), _keys(rng, 60_000, 200_000)
    va, vb = rng.integers(-99, 99, len(ka)), rng.integers(-99, 99, len(kb))
    vla, vlb = rng.random(len(ka)) > 0.2, rng.random(len(kb)) > 0.1
    _check(compile_op(mq.ojoin()), ka, va, kb, vb, vla, vlb)
    out = compile_op(mq.ojoin_even())(DeviceTable.from_columns([ka, va]),
                                      DeviceTable.from_columns([kb, vb]))
    t, d = merge.ojoin(ka, va, kb, vb)
    even = t % 2 == 0
    assert np.array_equal(out.column(0), t[even])
    assert np.array_equal(out.column(1), d[even])


def test_outer_join_float_values(device):
    rng = np.random.default_rng(6)
    ka, kb = _keys(rng, 30_000, 90_000), _keys(rng, 20_000, 90_000)
    va, vb = rng.standard_normal(len(ka)), rng.standard_normal(len(kb))
    out = compile_op(mq.ojoin(mq.MERGE_F))(DeviceTable.from_columns([ka, va]),
                                           DeviceTable.from_columns([kb, vb]))
    t, d = merge.ojoin(ka, va, kb, vb)
    assert np.array_equal(out.column(0), t)
    assert np.array_equal(out.column(1).view(np.int64), d.view(np.int64))


def test_outer_join_full_size(device):
    """20M + 16M keys, about half of B's keys also in A."""
    rng = np.random.default_rng(7)
    ka = np.sort(rng.choice(40_000_000, 20_000_000, replace=False)).astype(np.int64)
    kb = np.sort(np.concatenate([rng.choice(ka, 8_000_000, replace=False),
                                 rng.choice(np.arange(40_000_000, 80_000_000), 8_000_000,
                                            replace=False)])).astype(np.int64)
    kb = np.unique(kb)
    va, vb = rng.integers(-10**9, 10**9, len(ka)), rng.integers(-10**9, 10**9, len(kb))
    _check(compile_op(mq.ojoin()), ka, va, kb, vb)


@pytest.mark.parametrize("bad", ["unsorted", "duplicate"])
def test_outer_join_refuses_keys_not_strictly_ascending(device, bad):
    ka = np.arange(0, 1000, dtype=np.int64)
    if bad == "unsorted":
        ka[[10, 11]] = ka[[11, 10]]
    else:
        ka[500] = ka[499]
    kb = np.arange(0, 2000, 2, dtype=np.int64)
    q = compile_op(mq.ojoin())
    with pytest.raises(ReffineError, match="strictly ascending"):
        q(DeviceTable.from_columns([ka, ka]), DeviceTable.from_columns([kb, kb]))
    # the query still runs on good input afterwards
    _check(q, np.arange(5, dtype=np.int64), np.arange(5, dtype=np.int64), kb, kb)


@pytest.mark.parametrize("na,nb,oa,ob", [(0, 7, 0, -5), (9, 0, 3, 0), (100, 37, 0, -50),
                                         (4097, 5000, 12, -7), (100_000, 80_000, -1000, 250)])
def test_shifted_sides_against_oracle(device, na, nb, oa, ob):
    """_in(t + c, X) sides (ShiftedSpace): the merge runs on the keys plus
    their side's constant; outputs read each side's current row."""
    rng = np.random.default_rng(na + 3 * nb)
    span = max(2 * (na + nb), 4)
    ka, kb = _keys(rng, na, span), _keys(rng, nb, span)
    va = rng.integers(-10**12, 10**12, na)
    vb = rng.integers(-10**12, 10**12, nb)
    out = compile_op(mq.shifted(oa, ob))(DeviceTable.from_columns([ka, va]),
                                         DeviceTable.from_columns([kb, vb]))
    t, i, j = merge.union_states_np(ka, kb, off_a=oa, off_b=ob)
    assert np.array_equal(out.column(0), t)
    assert np.array_equal(out.column(1), merge.read_at(va, i) - merge.read_at(vb, j))


def test_algotrading_against_oracle(device):
    """benchmark/include/misc.h:15-61 as written: one stock_price table
    (ts = 0..n-1, val float32 ~ 100 + 10 N(0,1), the shape of
    benchmark/python/pandasbench.py's StockPrice) passed as both inputs, the
    second shifted by -50; t and the float32 differences bit for bit."""
    from reffine_amd import workloads
    n = 1_000_000
    rng = np.random.default_rng(21)
    ts = np.arange(n, dtype=np.int64)
    val = (rng.standard_normal(n) * 10 + 100).astype(np.float32)
    sp = DeviceTable.from_columns([ts, val])
    out = compile_op(workloads.algotrading())(sp, sp)
    t, d = merge.algotrading(ts, val, ts, val)
    assert out.nrows == n + 50
    assert np.array_equal(out.column(0), t)
    assert np.array_equal(out.column(1).view(np.uint32), d.view(np.uint32))


def test_shift_that_wraps_the_key_order_is_refused(device):
    """Keys + c wrap in int64 arithmetic; a wrap that breaks the ascending
    order is refused like unsorted keys."""
    big = np.iinfo(np.int64).max
    ka = np.array([big - 30, big - 5], np.int64)
    kb = np.array([0, 1], np.int64)
    q = compile_op(mq.shifted(20, 0))
    with pytest.raises(ReffineError, match="strictly ascending"):
        q(DeviceTable.from_columns([ka, ka]), DeviceTable.from_columns([kb, kb]))
