"""Pins the oracle before it is trusted: the CPU JIT restatement (oracle.cjit)
and the brute-force restatement (oracle.brute) against the reference test
suite's golden values and against closed forms / pure-Python loops; the
fixed-order restatement (oracle.detorder) against exact sums."""
import math

import numpy as np
import pytest

from oracle import brute, cjit, detorder, merge
from reffine_amd import datagen, workloads
import ref_queries as rq


def students_input():
    return {"cols": rq.students()}


@pytest.mark.parametrize("vectorize", [False, True])
def test_cjit_reduce_golden(vectorize):
    exp = rq.expected()
    (s,) = cjit.compile_op(rq.sum_hours_subvec(), vectorize)(students_input())
    assert s == (exp["aggregate_loop_test"],)
    (st,) = cjit.compile_op(rq.vector_op(), vectorize)(students_input())
    assert st[2] == exp["aggregate_op_test"]  # res = get(get(sum, 1), 0)


def test_cjit_transform_golden():
    exp = rq.expected()
    (cols,) = cjit.compile_op(rq.transform_op())(students_input())
    ids, hours = rq.students()[0], rq.students()[1]
    assert len(cols[0]) == exp["transform_op_test_len"]
    np.testing.assert_array_equal(cols[0], ids[5:])
    np.testing.assert_array_equal(cols[1], hours[5:] + 10)


def test_cjit_gen_table_and_join_golden():
    exp = rq.expected()
    gen = cjit.compile_op(rq.gen_table_op())
    tabs = []
    for lb, ub in [(0, 10), (5, 15), (7, 12)]:
        (cols,) = gen(lb, ub, cap=ub - lb)
        np.testing.assert_array_equal(cols[0], np.arange(lb, ub))
        tabs.append({"cols": [cols[0], cols[1]]})
    (out,) = cjit.compile_op(rq.join_op())(*tabs)
    np.testing.assert_array_equal(out[0], exp["join_op_test_keys"])
    np.testing.assert_array_equal(out[1], 3 * np.array(exp["join_op_test_keys"]))


def test_brute_scan_against_python_loops():
    n = 3000
    c = datagen.scan_table(n, 77)
    cnt, s = brute.scan_c1(c["k"], c["v"], c["d"])
    pc, ps = 0, []
    for k, v, d in zip(c["k"].tolist(), c["v"].tolist(), c["d"].tolist()):
        if 365 <= d < 730 and v < 0.5:
            pc += 1
            ps.append(v * float(k))
    assert cnt == pc and s == math.fsum(ps)
    cnt2, sq, svq, sel = brute.scan_c2(c["v"], c["d"], c["q"], d0=0, vlo=0.0, vhi=0.5, qmax=40)
    idx = [i for i in range(n) if 0 <= c["d"][i] < 365 and 0.0 <= c["v"][i] <= 0.5 and
           c["q"][i] < 40]
    assert cnt2 == len(idx) and list(sel) == idx
    assert sq == sum(int(c["q"][i]) for i in idx)


def test_cjit_matches_brute_on_workloads():
    n = 50_000
    c = datagen.scan_table(n, 5, with_q=True, rowid_dtype=np.int32)
    run = cjit.c2_query(d0=100, vlo=0.0, vhi=0.3, qmax=30)
    cnt, sq, svq, sel = run(c["rowid"], c["v"], c["d"], c["q"])
    w = brute.scan_c2(c["v"], c["d"], c["q"], d0=100, vlo=0.0, vhi=0.3, qmax=30, rowid=c["rowid"])
    assert (cnt, sq) == w[:2] and brute.rel_err(svq, w[2]) <= 1e-13
    np.testing.assert_array_equal(sel, w[3])
    c1 = datagen.scan_table(n, 6, with_q=False)
    (st,) = cjit.compile_op(workloads.scan_agg_c1())({"cols": [c1["rowid"], c1["k"], c1["v"], c1["d"]]})
    w1 = brute.scan_c1(c1["k"], c1["v"], c1["d"])
    assert st[0] == w1[0] and brute.rel_err(st[1], w1[1]) <= 1e-13


def test_cjit_prints_float_constants_like_cemitter():
    # std::to_string keeps 6 decimals (src/pass/cemitter.cpp:101).
    from reffine_amd.reffiner import _f64_t
    assert cjit.cpp_const(_f64_t, 0.05) == "double(0.050000)"
    assert cjit.cpp_const(_f64_t, 1e-7) == "double(0.000000)"


def test_detorder_integers_exact_and_fp_close():
    rng = np.random.default_rng(0)
    n = 100_000
    x = rng.integers(0, 1000, n).astype(np.int64)
    p = rng.random(n) < 0.3
    assert int(detorder.det_sum(x, p, 0, n, 256, 4, 4)) == int(x[p].sum())
    xf = rng.random(n)
    got = float(detorder.det_sum(xf, p, 0, n, 256, 4, 4))
    assert brute.rel_err(got, brute.fsum(xf[p])) < 1e-14
    # a deliberately wrong order (reversed rows) must not be bit-identical
    # for typical data, while a single-element input is exact in any order
    one = np.zeros(n)
    one[123] = 0.1
    assert float(detorder.det_sum(one, np.ones(n, bool), 0, n, 256, 4, 4)) == 0.1
    # unaligned start and tail handling
    got = float(detorder.det_sum(xf, p, 3, n - 5, 256, 4, 4))
    assert brute.rel_err(got, brute.fsum(xf[3:n - 5][p[3:n - 5]])) < 1e-14


def test_detorder_minmax_values_and_tie_order():
    """MAX / MIN in the fixed order: the value is the max / min over the
    selected non-NaN rows (std::max(s, x) with the state first drops a NaN
    row); among equal values (both signs of zero) the one that wins is the
    first row by (wave-tile, lane, row) when the tiles fit one block -- each
    combine keeps its left operand on a tie and the left operand holds the
    lower tiles, lanes and rows."""
    rng = np.random.default_rng(3)
    n = 50_000
    x = rng.normal(size=n)
    x[rng.random(n) < 0.1] = np.nan
    p = rng.random(n) < 0.4
    xs = x[p]
    assert detorder.det_fold(x, p, 0, n, 256, 4, 4, "max") == np.nanmax(xs)
    assert detorder.det_fold(x, p, 0, n, 256, 4, 4, "min") == np.nanmin(xs)
    assert detorder.det_fold(x, p, 5, n - 7, 256, 2, 8, "max") == np.nanmax(
        x[5:n - 7][p[5:n - 7]])
    # Ties: all selected values are zeros of random sign.
    block, items, sub = 256, 2, 4
    tile = sub * 64 * items
    for seed in range(6):
        r = np.random.default_rng(seed)
        m = 40 * tile + 77  # 41 tiles, fewer than the block's threads
        z = np.where(r.random(m) < 0.5, 0.0, -0.0)
        pz = r.random(m) < 0.01
        if not pz.any():
            continue
        rows = np.nonzero(pz)[0]
        tile_of = rows // tile
        lane_of = (rows % tile) // items % 64
        first = rows[np.lexsort((rows, lane_of, tile_of))[0]]
        for op in ("max", "min"):
            got = detorder.det_fold(z, pz, 0, m, block, items, sub, op)
            assert np.signbit(got) == np.signbit(z[first])


def test_q6_batch_oracle_against_loops():
    """brute.q6_batch against a row loop (every bound checked one row at a
    time), and a planted mistake (an exclusive disc bound) is caught."""
    c = datagen.group_table(3000, 77)
    d0, dlo, dhi, qmax = 700, 0.03, 0.05, 30
    cnt, sq, spd, rows = brute.q6_batch(c, d0, dlo, dhi, qmax)
    want = []
    for i in range(3000):
        d, disc, q = int(c["d"][i]), float(c["disc"][i]), int(c["qty"][i])
        if d0 <= d < d0 + 365 and dlo <= disc <= dhi and q < qmax:
            want.append(i)
    assert cnt == len(want) > 10
    assert rows.tolist() == want
    assert sq == sum(int(c["qty"][i]) for i in want)
    assert brute.rel_err(spd, math.fsum(float(c["price"][i]) * float(c["disc"][i])
                                        for i in want)) <= 1e-15
    wrong = [i for i in want if float(c["disc"][i]) < dhi]
    assert len(wrong) != cnt


def test_cjit_sequential_folds_against_row_loops():
    """The CPU JIT restatement folds the non-associative folds of
    tests/seq_queries.py exactly as plain Python row loops do (bit for bit),
    with and without null rows; a planted mistake (the decayed sum with the
    operands of + swapped, equal in IEEE arithmetic) must agree, one with the
    order of the last two rows swapped must not."""
    import seq_queries as sq
    rng = np.random.default_rng(8)
    n = 3000
    k = rng.integers(-1000, 1000, n).astype(np.int64)
    v = rng.random(n)
    for name, build in sq.BUILDERS.items():
        vv = v.copy()
        if name == "maxv":
            vv[rng.random(n) < 0.05] = np.nan
        (got,) = cjit.compile_op(build())({"cols": [np.arange(n), k, vv]})
        want = sq.loop(name, k, vv)
        assert [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in got] == \
            [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in want], name
    valid = rng.random(n) > 0.3
    (got,) = cjit.compile_op(sq.decay())({"cols": [np.arange(n), k, v],
                                         "valid": [None, valid, None]})
    assert got == sq.loop("decay", k, v, valid)
    vs = v.copy()
    vs[[n - 2, n - 1]] = vs[[n - 1, n - 2]]
    assert sq.loop("decay", k, vs) != sq.loop("decay", k, v)


def test_datagen_is_a_pure_function_of_seed_col_row():
    a = datagen.uniform(3, 2, 1000)
    b = datagen.uniform(3, 2, 600, start=400)
    np.testing.assert_array_equal(a[400:], b)
    assert not np.array_equal(a, datagen.uniform(4, 2, 1000))
    k = datagen.randint(1, 1, 100_000, 0, 1000, np.int64)
    assert k.min() == 0 and k.max() == 999
    g = datagen.group_table(100_000)
    f1 = np.bincount(g["f1"]) / 100_000
    assert abs(f1[0] - 0.25) < 0.01 and abs(f1[2] - 0.5) < 0.01


def test_group_join_oracle_against_nested_loops():
    """The brute-force group join (sorted keys, searchsorted, sequential
    np.add.at) equals the reference form's nested loops bit for bit,
    including groups without joined rows and unsorted build keys."""
    for seed, nb, n in ((1, 40, 3000), (2, 200, 5000), (3, 1, 100)):
        b, p = datagen.join_tables(nb, n, seed=seed)
        keys, sums, counts, _, _ = brute.group_join_c4(b, p)
        want = brute.group_join_nested(b, p)
        got = [(int(k), float(s), int(c)) for k, s, c in zip(keys, sums, counts)]
        assert got == want
    # a planted mistake (pairwise instead of sequential order) is caught
    b, p = datagen.join_tables(50, 20000, seed=4)
    keys, sums, counts, ranks, vals = brute.group_join_c4(b, p)
    pair = np.array([math.fsum([0.0]) if not (ranks == g).any() else
                     np.sum(vals[ranks == g]) for g in range(len(keys))])
    assert not np.array_equal(pair.view(np.int64), sums.view(np.int64))
    assert [(int(k), float(s), int(c)) for k, s, c in zip(keys, sums, counts)] == \
        brute.group_join_nested(b, p)


def test_group_fsums_are_correctly_rounded():
    ranks = np.array([0, 1, 0, 1, 0], dtype=np.int64)
    vals = np.array([1e16, 1.0, 1.0, 2.0, -1e16])
    s = brute.group_fsums(3, ranks, vals)
    assert list(s) == [1.0, 3.0, 0.0]


def test_cjit_grouped_aggregate_matches_brute():
    """The CPU JIT restatement of the Q1-idiom grouped aggregate (used as
    the grouped workload's CPU baseline) agrees with the brute force."""
    c = datagen.group_table(40_000, 3)
    fn = cjit.compile_op(workloads.group_agg_c3())
    (st,) = fn({"cols": [c["rowid"], c["f1"], c["f2"], c["qty"], c["price"], c["disc"],
                         c["tax"], c["d"]]})
    want = brute.group_c3(c)
    flat = list(st)
    assert len(flat) == len(want)
    for g, w in zip(flat, want):
        if isinstance(w, int):
            assert g == w
        else:
            assert brute.rel_err(g, w) <= 1e-12


def test_dim_join_oracle_against_loops():
    c = datagen.group_table(3000, 9)
    d = datagen.dim_table()
    keys, sums, counts = brute.dim_join(d, c, 2, 25)
    want = {}
    for k in d["dk"][d["dflag"] == 2]:
        want[int(k)] = [0.0, 0]
    for i in range(3000):
        k = int(c["d"][i])
        if c["qty"][i] < 25 and k in want:
            want[k][0] = want[k][0] + float(c["price"][i]) * (1.0 - float(c["disc"][i]))
            want[k][1] += 1
    assert list(keys) == sorted(want)
    assert [want[k][1] for k in sorted(want)] == list(counts)
    assert [want[k][0] for k in sorted(want)] == list(sums)


def _tpch_frames(t):
    import pandas as pd
    return (pd.DataFrame(t["lineitem"]), pd.DataFrame(t["orders"]),
            pd.DataFrame(t["customer"]))


def test_tpch_oracle_against_pandas():
    """oracle.tpch (row loops over a merge of the two key sets) against an
    independent group-by / join formulation of the same four queries in
    pandas: the same keys, exact counts and integer-valued sums, revenue
    within 1e-12 (pandas sums in its own order)."""
    from oracle import tpch as ot
    t = datagen.tpch_tables(6000, seed=5)
    li, od, cu = _tpch_frames(t)
    both = od[od.o_orderkey.isin(li.l_orderkey)]
    # Q18 (quantity 150 so that some orders qualify)
    tot = li.groupby("l_orderkey").l_quantity.sum()
    m = both.assign(tot=both.o_orderkey.map(tot))
    m = m[m.tot > 150].sort_values("o_orderkey")
    k, v = ot.q18(t, quantity=150)
    np.testing.assert_array_equal(k, m.o_orderkey.to_numpy())
    np.testing.assert_array_equal(v, m.tot.to_numpy())
    assert len(k) > 20
    # Q3
    date, seg = 795484800, 1
    rev = (li.l_extendedprice * (1.0 - li.l_discount)).where(li.l_shipdate > date, 0.0)
    r = rev.groupby(li.l_orderkey).sum()
    m = both.assign(rev=both.o_orderkey.map(r))
    m = m.merge(cu[["c_custkey", "c_mktsegment"]], left_on="o_custkey", right_on="c_custkey",
                how="inner")
    m = m[(m.o_orderdate < date) & (m.c_mktsegment == seg) & (m.rev > 0)]
    m = m.sort_values("o_orderkey")
    k, v = ot.q3(t, segment=seg, date=date)
    np.testing.assert_array_equal(k, m.o_orderkey.to_numpy())
    np.testing.assert_allclose(v, m.rev.to_numpy(), rtol=1e-12)
    assert len(k) > 10
    # Q4
    ex = (li.l_commitdate < li.l_receiptdate).groupby(li.l_orderkey).any()
    m = both[(both.o_orderdate >= 700000000) & (both.o_orderdate < 900000000)
             & both.o_orderkey.map(ex)]
    want = [int((m.o_orderpriority == p).sum()) for p in range(5)]
    assert ot.q4(t) == want
    # Q12
    k, h, lo = ot.q12(t)
    wk, wh, wl = _pandas_q12(t)
    np.testing.assert_array_equal(k, wk)
    np.testing.assert_array_equal(h, wh)
    np.testing.assert_array_equal(lo, wl)


def _pandas_q12(t):
    li, od, _ = _tpch_frames(t)
    both = od[od.o_orderkey.isin(li.l_orderkey)]
    ok = ((li.l_commitdate < li.l_receiptdate) & (li.l_shipdate < li.l_commitdate))
    n = ok.groupby(li.l_orderkey).sum()
    m = both.sort_values("o_orderkey")
    cnt = m.o_orderkey.map(n).to_numpy()
    high = np.isin(m.o_orderpriority.to_numpy(), [0, 1])
    return m.o_orderkey.to_numpy(), np.where(high, cnt, 0), np.where(high, 0, cnt)


def test_tpch_oracle_catches_a_planted_mistake():
    """The pandas formulation disagrees with an oracle that reads the ship
    date where Q12 reads the commit date (the mistake planted by running
    the oracle on lineitem columns swapped accordingly)."""
    from oracle import tpch as ot
    t = datagen.tpch_tables(3000, seed=9)
    bad = {**t, "lineitem": {**t["lineitem"], "l_commitdate": t["lineitem"]["l_shipdate"]}}
    assert not np.array_equal(ot.q12(bad)[1], _pandas_q12(t)[1])


def test_tpch_vectorised_oracle_matches_the_row_loops():
    """oracle.tpch's numpy forms (used where the row loops are too slow)
    give the row loops' results bit for bit."""
    from oracle import tpch as ot
    t = datagen.tpch_tables(8000, seed=21)
    for loop, vec, a in ((ot.q18, ot.q18_np, (120,)), (ot.q3, ot.q3_np, ()),
                         (ot.q12, ot.q12_np, ())):
        for x, y in zip(loop(t, *a), vec(t, *a)):
            np.testing.assert_array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    assert ot.q4(t) == ot.q4_np(t)


def test_tpch_example_oracle_against_pandas():
    """oracle.tpch.example (row loops) against a pandas merge + group-by, and
    its numpy form against the loops."""
    import pandas as pd
    from oracle import tpch as ot
    t = datagen.tpch_parts(4000, seed=8)
    ps, part, sup = (pd.DataFrame(t["partsupp"]), pd.DataFrame(t["part"]),
                     pd.DataFrame(t["supplier"]))
    for nation, size in ((0, 15), (7, 30)):
        m = ps.merge(sup[["s_suppkey", "s_nationkey"]], left_on="ps_suppkey", right_on="s_suppkey")
        m["q"] = np.where(m.s_nationkey == nation, m.ps_availqty, 0)
        tot = m.groupby("ps_partkey").q.sum()
        pp = part[part.p_size > size].sort_values("p_partkey")
        pp = pp[pp.p_partkey.isin(tot.index)]
        k, v = ot.example(t, nation, size)
        np.testing.assert_array_equal(k, pp.p_partkey.to_numpy())
        np.testing.assert_array_equal(v, tot.loc[pp.p_partkey].to_numpy())
        kn, vn = ot.example_np(t, nation, size)
        np.testing.assert_array_equal(k, kn)
        np.testing.assert_array_equal(v, vn)


def test_tpch_fold_oracles_against_loops():
    """oracle.tpch's Q1 / Q6 / TPC-DS Q9 (numpy masks + fsum) against plain
    row loops with math.fsum on small tables, and Q11 (row loops) against a
    pandas merge + group-by."""
    import math
    import pandas as pd
    from oracle import tpch as ot
    t = datagen.tpch_tables(1500, seed=13)
    li = t["lineitem"]
    n = len(np.unique(li["l_orderkey"]))  # runs of the sorted key column
    acc = [[] for _ in range(15)]
    for r in range(n):
        g = int(li["l_returnflag"][r])
        ep, disc, tax = li["l_extendedprice"][r], li["l_discount"][r], li["l_tax"][r]
        dp = ep * (1.0 - disc)
        for j, x in enumerate((li["l_quantity"][r], ep, dp, dp * (1.0 + tax), 1.0)):
            acc[5 * g + j].append(x)
    assert ot.q1(t) == [math.fsum(a) for a in acc]
    s6 = [li["l_extendedprice"][r] * li["l_discount"][r] for r in range(len(li["l_orderkey"]))
          if 820454400 <= li["l_shipdate"][r] < 852076800 and 0.04 <= li["l_discount"][r] <= 0.06
          and li["l_quantity"][r] < 24.5]
    assert s6 and ot.q6(t) == math.fsum(s6)
    ss = datagen.store_sales(3000, seed=3)
    for b, (c, a, p) in enumerate(ot.tpcds_q9(ss)):
        rows = [r for r in range(3000) if 20 * b + 1 <= ss["ss_quantity"][r] <= 20 * b + 20]
        assert c == len(rows)
        assert a == math.fsum(ss["ss_ext_tax"][r] for r in rows)
        assert p == math.fsum(ss["ss_net_paid_inc_tax"][r] for r in rows)
    tp = datagen.tpch_parts(3000, seed=2)
    ps, sup = pd.DataFrame(tp["partsupp"]), pd.DataFrame(tp["supplier"])
    m = ps.merge(sup[["s_suppkey", "s_nationkey"]], left_on="ps_suppkey", right_on="s_suppkey")
    m["v"] = np.where(m.s_nationkey == 4, m.ps_availqty.astype(np.float64) * m.ps_supplycost, 0.0)
    tot = m.groupby("ps_partkey").v.sum()
    tot = tot[tot > 0]
    k, v = ot.q11(tp, 4)
    np.testing.assert_array_equal(k, tot.index.to_numpy())
    np.testing.assert_allclose(v, tot.to_numpy(), rtol=1e-12)


def test_pagerank_oracle_against_sparse_matrix():
    """oracle.graph.pagerank against the matrix form of one PageRank step:
    new = alpha * A^T (pr / outdeg) + (1 - alpha) / n over the nodes with
    in-edges (float64, within float32 rounding), and a planted mistake (the
    edge multiplicity dropped) that the check catches."""
    import scipy.sparse as sp
    from oracle import graph as og
    n = 800
    t = datagen.graph_tables(n, 6000, seed=9)
    e = t["edges"]
    A = sp.coo_matrix((e["count"].astype(np.float64), (e["src"], e["dst"])), shape=(n, n)).tocsr()
    deg = np.asarray(A.sum(axis=1)).ravel()
    pr = t["pr"]["rank"].astype(np.float64)
    contrib = np.divide(pr, deg, out=np.zeros(n), where=deg > 0)
    want = 0.85 * (A.T @ contrib) + 0.15 / n
    k, v = og.pagerank(t, n)
    np.testing.assert_allclose(v, want[k], rtol=1e-5)
    assert set(k.tolist()) == set(np.flatnonzero(np.asarray(A.sum(axis=0)).ravel() > 0).tolist())
    bad = {**t, "rev_edges": {**t["rev_edges"], "count": np.ones_like(t["rev_edges"]["count"])}}
    if (t["rev_edges"]["count"] > 1).any():
        assert not np.allclose(og.pagerank(bad, n)[1], want[k], rtol=1e-5)


def test_graph_vectorised_oracles_match_the_loops():
    """oracle.graph's numpy forms equal its row loops bit for bit."""
    from oracle import graph as og
    t = datagen.graph_tables(700, 5000, seed=4)
    k, v = og.pagerank(t, 700)
    kn, vn = og.pagerank_np(t, 700)
    np.testing.assert_array_equal(k, kn)
    np.testing.assert_array_equal(v.view(np.uint32), vn.view(np.uint32))
    b = datagen.bodies(48, seed=2)
    for x, y in zip(og.nbody(b, 48), og.nbody_np(b, 48)):
        np.testing.assert_array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))


def _subsets(universe):
    for mask in range(1 << len(universe)):
        yield np.array([u for b, u in enumerate(universe) if mask >> b & 1], np.int64)


def test_union_merge_closed_form_against_the_loop_by_brute_force():
    """oracle/merge.py's vectorised states (sorted union of the keys, rows by
    searchsorted) equal its row-by-row merge loop on every pair of subsets of
    a 5-key universe, with random key validity; a loop that advances only A
    on a tie (emitting the key twice) must disagree somewhere."""
    rng = np.random.default_rng(11)
    subsets = list(_subsets([-3, 0, 2, 7, 9]))
    for ka in subsets:
        for kb in subsets:
            for trial in range(2):
                va = rng.random(len(ka)) > 0.3 if trial else None
                vb = rng.random(len(kb)) > 0.3 if trial else None
                loop = merge.union_states_loop(ka, kb, va, vb)
                t, i, j = merge.union_states_np(ka, kb, va, vb)
                assert loop == list(zip(t.tolist(), i.tolist(), j.tolist())), (ka, kb, va, vb)

    def tie_advances_a_only(ka, kb):
        out, i, j = [], 0, 0
        while i < len(ka) or j < len(kb):
            x = ka[i] if i < len(ka) else None
            y = kb[j] if j < len(kb) else None
            t = x if y is None or (x is not None and x <= y) else y
            out.append((int(t), i, j))
            if x is not None and (y is None or x <= y):
                i += 1
            else:
                j += 1
        return out
    bad = 0
    for ka in subsets:
        for kb in subsets:
            t, i, j = merge.union_states_np(ka, kb)
            bad += tie_advances_a_only(ka, kb) != list(zip(t.tolist(), i.tolist(), j.tolist()))
    assert bad > 0


def test_shifted_union_merge_against_the_loop_and_a_closed_form():
    """Shifted sides (_in(t + c, X): the loop runs on X's keys + c) -- the
    vectorised states equal the row loop for every pair of subsets of a
    4-key universe and offsets on both sides; and AlgoTrading's states over
    keys 0..n-1 on both sides with B shifted by -50 follow from hand
    counting: t = -50..-1 at A row 0 and B row t + 50, t = 0..n-51 at rows
    t and t + 50, t = n-50..n-1 at A row t with B run out (B reads 0)."""
    subsets = list(_subsets([-2, 0, 3, 5]))
    for oa, ob in [(0, -3), (2, 0), (-1, 4), (7, 7)]:
        for ka in subsets:
            for kb in subsets:
                loop = merge.union_states_loop(ka, kb, off_a=oa, off_b=ob)
                t, i, j = merge.union_states_np(ka, kb, off_a=oa, off_b=ob)
                assert loop == list(zip(t.tolist(), i.tolist(), j.tolist())), (ka, kb, oa, ob)
    n = 200
    k = np.arange(n, dtype=np.int64)
    t, i, j = merge.union_states_np(k, k, off_b=-50)
    assert t.tolist() == list(range(-50, n))
    assert i.tolist() == [0] * 50 + list(range(n))
    assert j.tolist() == list(range(50)) + [r + 50 for r in range(n - 50)] + [n] * 50
    v = np.linspace(90, 110, n).astype(np.float32)
    t2, d = merge.algotrading(k, v, k, v)
    want = np.concatenate([v[0] - v[:50], v[:n - 50] - v[50:], v[n - 50:]]).astype(np.float32)
    assert np.array_equal(t2, t) and np.array_equal(d.view(np.uint32), want.view(np.uint32))


def _reference_loop_defined_prefix(ka, kb):
    """The reference's loop transcribed literally (locate through a first-
    occurrence index, min of the current keys, lc / rc advance), stopped at
    the first read outside either vector: the states it defines."""
    ia = {}
    for r, k in enumerate(ka.tolist()):
        ia.setdefault(k, r)
    ib = {}
    for r, k in enumerate(kb.tolist()):
        ib.setdefault(k, r)
    if not len(ka) or not len(kb):
        return []
    lb = min(int(ka[0]), int(kb[0]))
    i, j = ia.get(lb, -1), ib.get(lb, -1)
    out = []
    while i < len(ka) or j < len(kb):
        if not (0 <= i < len(ka) and 0 <= j < len(kb)):
            break
        t = min(int(ka[i]), int(kb[j]))
        out.append((t, i, j))
        lc, rc = ka[i] <= kb[j], kb[j] <= ka[i]
        i, j = i + lc, j + rc
    return out


def test_union_merge_matches_the_reference_loop_where_it_is_defined():
    """Where the reference's loop stays inside both vectors (equal first
    keys), its states are the first states of oracle/merge.py's loop; and the
    micro.h case (the same table on both sides) emits every key once with
    diff 0."""
    rng = np.random.default_rng(12)
    for _ in range(300):
        u = np.sort(rng.choice(60, size=30, replace=False)).astype(np.int64)
        ka = np.sort(np.unique(np.concatenate([[u[0]], rng.choice(u, 12)])))
        kb = np.sort(np.unique(np.concatenate([[u[0]], rng.choice(u, 12)])))
        ref = _reference_loop_defined_prefix(ka, kb)
        ours = merge.union_states_loop(ka, kb)
        assert len(ref) >= 1 and ours[:len(ref)] == ref
    k = np.arange(0, 50, 3, dtype=np.int64)
    v = rng.integers(-9, 9, len(k))
    t, d = merge.ojoin(k, v, k, v)
    assert t.tolist() == k.tolist() and not d.any()


def _bits(vals):
    return [np.float64(x).view(np.uint64) if isinstance(x, float) else x for x in vals]


def test_cjit_fixed_order_equals_detorder_on_workloads():
    """oracle/cjit.py's fixed-order emission (C text in the GPU deterministic
    mode's tile / butterfly / completion order, compiled by clang) equals the
    numpy restatement of that order bit for bit on the C1, C2 and C3 folds,
    for unaligned sizes and a few launch shapes; on contributions where the
    order shows, two shapes give different bits, each matching its own numpy
    order, so the order is what is pinned."""
    n = 70_001
    c1 = datagen.scan_table(n, 6, with_q=False)
    m1 = (c1["d"] >= 365) & (c1["d"] < 730) & (c1["v"] < 0.5)
    x1 = c1["v"] * c1["k"].astype(np.float64)
    for shape in [(256, 2, 8), (512, 4, 2), (256, 1, 1)]:
        (st,) = cjit.compile_op(workloads.scan_agg_c1(), fixed_order=shape)(
            {"cols": [c1["rowid"], c1["k"], c1["v"], c1["d"]]})
        assert st[0] == int(m1.sum())
        assert _bits([st[1]]) == _bits([float(detorder.det_sum(x1, m1, 0, n, *shape))]), shape
    # Contributions of both signs spread over 15 decades make the order visible.
    rng = np.random.default_rng(1)
    kw = (rng.choice([-1, 1], n) * (10.0 ** rng.uniform(0, 15, n))).astype(np.int64)
    xw = c1["v"] * kw.astype(np.float64)
    cols = {"cols": [c1["rowid"], kw, c1["v"], c1["d"]]}
    (a,) = cjit.compile_op(workloads.scan_agg_c1(), fixed_order=(256, 1, 1))(cols)
    (b,) = cjit.compile_op(workloads.scan_agg_c1(), fixed_order=(256, 2, 8))(cols)
    assert _bits([a[1]]) == _bits([float(detorder.det_sum(xw, m1, 0, n, 256, 1, 1))])
    assert _bits([b[1]]) == _bits([float(detorder.det_sum(xw, m1, 0, n, 256, 2, 8))])
    assert _bits([a[1]]) != _bits([b[1]])

    c2 = datagen.scan_table(n, 5, with_q=True, rowid_dtype=np.int32)
    cq = cjit.compile_op(workloads.scan_select_c2(), fixed_order=(256, 4, 4))
    (st2, sel) = cq({"cols": [c2["rowid"], c2["k"], c2["v"], c2["d"], c2["q"]]})
    m2 = (c2["d"] >= 1000) & (c2["d"] < 1365) & (c2["v"] >= 0.05) & (c2["v"] <= 0.07) & \
        (c2["q"] < 24)
    assert st2[0] == int(m2.sum()) and st2[1] == int(c2["q"][m2].sum())
    x2 = c2["v"] * c2["q"].astype(np.float64)
    assert _bits([st2[2]]) == _bits([float(detorder.det_sum(x2, m2, 0, n, 256, 4, 4))])
    np.testing.assert_array_equal(sel[0], c2["rowid"][m2])

    c3 = datagen.group_table(n, 3)
    (st3,) = cjit.compile_op(workloads.group_agg_c3(), fixed_order=(256, 2, 16))(
        {"cols": [c3[k] for k in ("rowid", "f1", "f2", "qty", "price", "disc", "tax", "d")]})
    keep = c3["d"] <= 2526
    vals = brute.group_c3_contribs(c3)
    want = []
    for a, b in workloads.C3_GROUPS:
        m = keep & (c3["f1"] == a) & (c3["f2"] == b)
        for x in vals + [np.ones(n, np.int64)]:
            want.append(detorder.det_sum(x, m, 0, n, 256, 2, 16).item())
    assert _bits(list(st3)) == _bits(want)


def _nbody_loops(b, n, g=1.0, dt=0.01, assoc="left", compact=True):
    """reffine's Nbody step (benchmark/include/misc.h:77-202) as plain
    Python loops over Python floats: the pair Op's nested loops (id_a outer,
    id_b inner, a == b skipped), each expression one IEEE operation in the
    C text's order, the per-body window sums over pair rows
    [idx*n, (idx+1)*n) that exist, then the velocity and position update.
    `assoc` and `compact` plant mistakes for the check below."""
    m = [float(v) for v in b["m"]]
    x, y, z = ([float(v) for v in b[c]] for c in ("x", "y", "z"))
    pairs = []
    for ia in range(n):
        for ib in range(n):
            if ia == ib and compact:
                continue
            dx, dy, dz = x[ia] - x[ib], y[ia] - y[ib], z[ia] - z[ib]
            if assoc == "left":
                dist2 = (dx * dx + dy * dy) + dz * dz
            else:
                dist2 = dx * dx + (dy * dy + dz * dz)
            if dist2 == 0.0:
                pairs.append((0.0, 0.0, 0.0))
                continue
            dist = math.sqrt(dist2)
            f = (g * (m[ia] * m[ib])) / dist2
            pairs.append(((f * dx) / dist, (f * dy) / dist, (f * dz) / dist))
    out = {c: [] for c in ("X", "Y", "Z", "VX", "VY", "VZ")}
    for idx in range(n):
        s = [0.0, 0.0, 0.0]
        for r in range(idx * n, (idx + 1) * n):
            if r < len(pairs):
                for c in range(3):
                    s[c] = s[c] + pairs[r][c]
        v = [float(b[k][idx]) + (s[c] / m[idx]) * dt for c, k in enumerate(("vx", "vy", "vz"))]
        p = [(x, y, z)[c][idx] + v[c] * dt for c in range(3)]
        for c, k in enumerate(("X", "Y", "Z")):
            out[k].append(p[c])
        for c, k in enumerate(("VX", "VY", "VZ")):
            out[k].append(v[c])
    return out, pairs


def test_nbody_oracle_against_python_loops():
    """oracle/graph.py's nbody (and its vectorised form) equal independent
    Python loops written from misc.h bit for bit on n <= 6 bodies; the pair
    forces are antisymmetric (F(a, b) = -F(b, a) exactly: IEEE subtraction
    is, and the rest is symmetric in a and b); a different association of
    dist2 or pair rows kept for a == b give other bits, so both are pinned."""
    from oracle import graph as og
    caught = {"right": 0, "nocompact": 0}
    for n in range(2, 7):
        for seed in range(4):
            b = datagen.bodies(n, seed=100 + seed)
            want, pairs = _nbody_loops(b, n)
            for got in (og.nbody(b, n), og.nbody_np(b, n)):
                for k, col in zip(("X", "Y", "Z", "VX", "VY", "VZ"), got[2:]):
                    assert np.asarray(col, np.float64).view(np.uint64).tolist() == \
                        np.asarray(want[k], np.float64).view(np.uint64).tolist(), (n, seed, k)
            rows = [(ia, ib) for ia in range(n) for ib in range(n) if ia != ib]
            at = {ab: f for ab, f in zip(rows, pairs)}
            for (ia, ib), f in at.items():
                assert all(u == -v for u, v in zip(f, at[(ib, ia)]))
            for plant, kw in (("right", {"assoc": "right"}), ("nocompact", {"compact": False})):
                bad, _ = _nbody_loops(b, n, **kw)
                caught[plant] += any(bad[k] != want[k] for k in want)
    assert caught["right"] > 0 and caught["nocompact"] > 0
