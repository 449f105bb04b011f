"""Seeded synthetic tables for the scan / group-by / join workloads.

Every value is a pure function of (seed, column, row): a SplitMix64
finaliser over a counter, so any row range can be regenerated alone and
the same table is produced on every machine. Uniform doubles are
(x >> 11) * 2^-53; uniform integers in [lo, hi) are lo + (x >> 11) % (hi-lo).
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_C = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, col: int, n: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _G + np.uint64(col) * _C
        rows = np.arange(start, start + n, dtype=np.uint64)
        return mix64(rows + base)


def uniform(seed: int, col: int, n: int, start: int = 0) -> np.ndarray:
    return (u64(seed, col, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def randint(seed: int, col: int, n: int, lo: int, hi: int, dtype, start: int = 0) -> np.ndarray:
    x = u64(seed, col, n, start) >> np.uint64(11)
    return (np.int64(lo) + (x % np.uint64(hi - lo)).astype(np.int64)).astype(dtype)


def scan_table(n: int, seed: int = 42, with_q: bool = True, rowid_dtype=np.int64):
    """Scan workload columns: rowid, k:i64 U[0,1000), v:f64 U[0,1),
    d:i32 U[0,2556) and (optionally) q:i64 U[1,50]."""
    cols = {
        "rowid": np.arange(n, dtype=rowid_dtype),
        "k": randint(seed, 1, n, 0, 1000, np.int64),
        "v": uniform(seed, 2, n),
        "d": randint(seed, 3, n, 0, 2556, np.int32),
    }
    if with_q:
        cols["q"] = randint(seed, 4, n, 1, 51, np.int64)
    return cols


def group_table(n: int, seed: int = 7):
    """Low-cardinality group-by workload: f1 in {0,1,2} (25/25/50),
    f2 in {0,1}, qty U[1,50], price = qty*U[900,105000]/100,
    disc in {0.00..0.10}, tax in {0.00..0.08}, d U[0,2556)."""
    r = randint(seed, 1, n, 0, 4, np.int64)
    f1 = np.where(r >= 2, 2, r).astype(np.int8)
    f2 = randint(seed, 2, n, 0, 2, np.int8)
    qty = randint(seed, 3, n, 1, 51, np.int64)
    price = qty.astype(np.float64) * randint(seed, 4, n, 900, 105001, np.int64).astype(np.float64) / 100.0
    disc = randint(seed, 5, n, 0, 11, np.int64).astype(np.float64) / 100.0
    tax = randint(seed, 6, n, 0, 9, np.int64).astype(np.float64) / 100.0
    d = randint(seed, 7, n, 0, 2556, np.int32)
    return {"rowid": np.arange(n, dtype=np.int64), "f1": f1, "f2": f2, "qty": qty,
            "price": price, "disc": disc, "tax": tax, "d": d}


def permutation(seed: int, col: int, n: int) -> np.ndarray:
    """Seeded permutation of [0, n): rows sorted by their counter hash."""
    return np.argsort(u64(seed, col, n), kind="stable").astype(np.int64)


def join_tables(n_build: int, n_probe: int, seed: int = 11):
    """Hash-join workload: build bk = permutation of [0,n_build), bd, bf;
    probe pk uniform over build keys, pv U[0,1), pq U[1,50]."""
    build = {"bk": permutation(seed, 1, n_build),
             "bd": randint(seed, 2, n_build, 0, 2556, np.int32),
             "bf": randint(seed, 3, n_build, 0, 5, np.int8)}
    probe = {"rowid": np.arange(n_probe, dtype=np.int64),
             "pk": randint(seed, 4, n_probe, 0, n_build, np.int64),
             "pv": uniform(seed, 5, n_probe),
             "pq": randint(seed, 6, n_probe, 1, 51, np.int64)}
    return build, probe


def lognormal_sizes(n: int, seed: int, median: int = 262_144, sigma: float = 1.2,
                    lo: int = 4096, hi: int = 8_388_608) -> np.ndarray:
    """Seeded log-normal row counts clipped to [lo, hi] (Box-Muller over two
    counter-based uniforms). With 64 batches and seed 5: 43.8M rows, median
    266K, from 15K to 7.9M."""
    u1 = uniform(seed, 1, n)
    u2 = uniform(seed, 2, n)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    return np.clip(np.rint(median * np.exp(sigma * z)), lo, hi).astype(np.int64)


def poisson_arrivals(n: int, qps: float, seed: int = 23) -> np.ndarray:
    """Arrival times in microseconds of a seeded Poisson process."""
    gaps = -np.log1p(-uniform(seed, 1, n)) / qps * 1e6
    return np.cumsum(gaps) - gaps[0]


def dim_table(seed: int = 31, days: int = 2556) -> dict:
    """Date dimension: dk = 0..days-1, dflag uniform in [0, 5)."""
    return {"dk": np.arange(days, dtype=np.int64), "dflag": randint(seed, 1, days, 0, 5, np.int8)}


_DAY = 86_400
_EPOCH_1992 = 694_224_000  # 1992-01-01 in seconds


def tpch_tables(n_orders: int, seed: int = 53) -> dict:
    """TPC-H-shaped lineitem / orders / customer columns (numeric codes in
    place of strings, dates in epoch seconds), laid out as reffine's TPC-H
    programs read them (benchmark/include/tpch.h): lineitem sorted by
    orderkey with 1..7 lines per order; orders with sparse keys
    (8 of every 32, as dbgen); customers n_orders/10.

    So that every join path is exercised: 1 order in 16 is missing from the
    orders table (its lines stay), 1 in 13 has no lines (orders only), and
    1 in 11 names a customer key past the customer table (Locate misses).
    """
    i = np.arange(n_orders, dtype=np.int64)
    okey = (i // 8) * 32 + (i % 8) + 1
    n_cust = max(1, n_orders // 10)
    odate = _EPOCH_1992 + randint(seed, 1, n_orders, 0, 2406, np.int64) * _DAY
    ocust = randint(seed, 2, n_orders, 1, n_cust + 1, np.int64)
    ocust = np.where(randint(seed, 3, n_orders, 0, 11, np.int64) == 0, n_cust + 1 + i, ocust)
    in_orders = randint(seed, 4, n_orders, 0, 16, np.int64) != 0
    nlines = randint(seed, 5, n_orders, 1, 8, np.int64)
    nlines = np.where(randint(seed, 6, n_orders, 0, 13, np.int64) == 0, 0, nlines)
    nl = int(nlines.sum())
    lo = np.repeat(np.arange(n_orders), nlines)          # order of each line
    line = np.arange(nl, dtype=np.int64) - np.repeat(np.cumsum(nlines) - nlines, nlines)
    qty = randint(seed, 7, nl, 1, 51, np.int64).astype(np.float64)
    pprice = randint(seed, 8, nl, 90_000, 200_001, np.int64).astype(np.float64) / 100.0
    ship = odate[lo] + randint(seed, 9, nl, 1, 122, np.int64) * _DAY
    lineitem = {
        "l_orderkey": okey[lo],
        "l_partkey": randint(seed, 10, nl, 1, 200_001, np.int64),
        "l_suppkey": randint(seed, 11, nl, 1, 10_001, np.int64),
        "l_linenumber": line + 1,
        "l_quantity": qty,
        "l_extendedprice": np.round(qty * pprice, 2),
        "l_discount": randint(seed, 12, nl, 0, 11, np.int64).astype(np.float64) / 100.0,
        "l_tax": randint(seed, 13, nl, 0, 9, np.int64).astype(np.float64) / 100.0,
        "l_returnflag": randint(seed, 14, nl, 0, 3, np.int32),
        "l_linestatus": randint(seed, 15, nl, 0, 2, np.int8),
        "l_shipdate": ship,
        "l_commitdate": odate[lo] + randint(seed, 16, nl, 30, 91, np.int64) * _DAY,
        "l_receiptdate": ship + randint(seed, 17, nl, 1, 31, np.int64) * _DAY,
        "l_shipinstruct": randint(seed, 18, nl, 0, 4, np.int8),
        "l_shipmode": randint(seed, 19, nl, 0, 7, np.int8),
        "l_comment": randint(seed, 20, nl, 0, 1 << 40, np.int64),
    }
    keep = in_orders
    orders = {
        "o_orderkey": okey[keep],
        "o_custkey": ocust[keep],
        "o_orderstatus": randint(seed, 21, n_orders, 0, 3, np.int8)[keep],
        "o_totalprice": (randint(seed, 22, n_orders, 100_000, 50_000_001, np.int64)
                         .astype(np.float64) / 100.0)[keep],
        "o_orderdate": odate[keep],
        "o_orderpriority": randint(seed, 23, n_orders, 0, 5, np.int8)[keep],
        "o_clerk": randint(seed, 24, n_orders, 1, 1001, np.int64)[keep],
        "o_shippriority": np.zeros(int(keep.sum()), dtype=np.int32),
        "o_comment": randint(seed, 25, n_orders, 0, 1 << 40, np.int64)[keep],
    }
    ci = np.arange(n_cust, dtype=np.int64)
    customer = {
        "c_custkey": ci + 1,
        "c_name": randint(seed, 26, n_cust, 0, 1 << 40, np.int64),
        "c_address": randint(seed, 27, n_cust, 0, 1 << 40, np.int64),
        "c_nationkey": randint(seed, 28, n_cust, 0, 25, np.int64),
        "c_phone": randint(seed, 29, n_cust, 0, 1 << 40, np.int64),
        "c_acctbal": (randint(seed, 30, n_cust, -99_999, 1_000_000, np.int64)
                      .astype(np.float64) / 100.0),
        "c_mktsegment": randint(seed, 31, n_cust, 0, 5, np.int8),
        "c_comment": randint(seed, 32, n_cust, 0, 1 << 40, np.int64),
    }
    return {"lineitem": lineitem, "orders": orders, "customer": customer}


def tpch_parts(n_parts: int, seed: int = 59) -> dict:
    """TPC-H-shaped part / supplier / partsupp columns as reffine's
    TPCHQueryExample reads them (benchmark/include/tpch.h:606-682): partsupp
    sorted by partkey with 4 suppliers per part (its partkey run-end
    encoded), suppliers n_parts/20; every partsupp key exists in part and
    supplier."""
    n_supp = max(1, n_parts // 20)
    pk = np.arange(1, n_parts + 1, dtype=np.int64)
    ps_part = np.repeat(pk, 4)
    j = np.tile(np.arange(4, dtype=np.int64), n_parts)
    # dbgen's supplier spread for a part: (p + j * (S/4 + (p-1)/S)) mod S + 1
    ps_supp = (ps_part + j * (n_supp // 4 + (ps_part - 1) // n_supp)) % n_supp + 1
    n_ps = len(ps_part)
    part = {
        "p_partkey": pk,
        "p_name": randint(seed, 1, n_parts, 0, 1 << 40, np.int64),
        "p_mfgr": randint(seed, 2, n_parts, 1, 6, np.int64),
        "p_brand": randint(seed, 3, n_parts, 1, 26, np.int64),
        "p_type": randint(seed, 4, n_parts, 0, 150, np.int64),
        "p_size": randint(seed, 5, n_parts, 1, 51, np.int32),
        "p_container": randint(seed, 6, n_parts, 0, 40, np.int64),
        "p_retailprice": (90_000 + (pk % 20_001) + 100 * (pk % 1000)).astype(np.float64) / 100.0,
        "p_comment": randint(seed, 7, n_parts, 0, 1 << 40, np.int64),
    }
    sk = np.arange(1, n_supp + 1, dtype=np.int64)
    supplier = {
        "s_suppkey": sk,
        "s_name": randint(seed, 8, n_supp, 0, 1 << 40, np.int64),
        "s_address": randint(seed, 9, n_supp, 0, 1 << 40, np.int64),
        "s_nationkey": randint(seed, 10, n_supp, 0, 25, np.int64),
        "s_phone": randint(seed, 11, n_supp, 0, 1 << 40, np.int64),
        "s_acctbal": (randint(seed, 12, n_supp, -99_999, 1_000_000, np.int64)
                      .astype(np.float64) / 100.0),
        "s_comment": randint(seed, 13, n_supp, 0, 1 << 40, np.int64),
    }
    partsupp = {
        "ps_partkey": ps_part,
        "ps_suppkey": ps_supp,
        "ps_availqty": randint(seed, 14, n_ps, 1, 10_000, np.int32),
        "ps_supplycost": (randint(seed, 15, n_ps, 100, 100_001, np.int64)
                          .astype(np.float64) / 100.0),
        "ps_comment": randint(seed, 16, n_ps, 0, 1 << 40, np.int64),
    }
    return {"part": part, "supplier": supplier, "partsupp": partsupp}


def store_sales(n: int, seed: int = 61) -> dict:
    """TPC-DS-shaped store_sales columns as reffine's TPCDSQuery9 reads them
    (benchmark/include/tpch.h:218-313): 23 columns in TPC-DS order, keys as
    int64, ss_quantity int32 in [1, 100], amounts float64 in cents."""
    names = ["ss_sold_date_sk", "ss_sold_time_sk", "ss_item_sk", "ss_customer_sk", "ss_cdemo_sk",
             "ss_hdemo_sk", "ss_addr_sk", "ss_store_sk", "ss_promo_sk", "ss_ticket_number",
             "ss_quantity", "ss_wholesale_cost", "ss_list_price", "ss_sales_price",
             "ss_ext_discount_amt", "ss_ext_sales_price", "ss_ext_wholesale_cost",
             "ss_ext_list_price", "ss_ext_tax", "ss_coupon_amt", "ss_net_paid",
             "ss_net_paid_inc_tax", "ss_net_profit"]
    out = {}
    for c, name in enumerate(names):
        if c < 10:
            out[name] = randint(seed, c + 1, n, 0, 1 << 30, np.int64)
        elif c == 10:
            out[name] = randint(seed, c + 1, n, 1, 101, np.int32)
        else:
            out[name] = randint(seed, c + 1, n, 0, 2_000_000, np.int64).astype(np.float64) / 100.0
    out["ss_sold_date_sk"] = np.sort(out["ss_sold_date_sk"])
    return out


def graph_tables(n_nodes: int, n_edges: int, seed: int = 67) -> dict:
    """A directed multigraph as reffine's PageRank reads it (benchmark/
    include/misc.h:212-291): edges (src, dst, count) sorted by (src, dst)
    and rev_edges (dst, src, count) sorted by (dst, src), each pair once
    with its multiplicity; pr (id, rank float32 = 1/n) for every node.
    About 1 node in 8 has no outgoing edge."""
    src = randint(seed, 1, n_edges, 0, n_nodes, np.int64)
    dst = randint(seed, 2, n_edges, 0, n_nodes, np.int64)
    keep = (src % 8) != 3  # nodes with no outgoing edges
    src, dst = src[keep], dst[keep]
    pair = np.unique(src * n_nodes + dst, return_counts=True)
    s, d, c = pair[0] // n_nodes, pair[0] % n_nodes, pair[1].astype(np.int64)
    r = np.lexsort((s, d))
    return {
        "edges": {"src": s, "dst": d, "count": c},
        "rev_edges": {"dst": d[r], "src": s[r], "count": c[r]},
        "pr": {"id": np.arange(n_nodes, dtype=np.int64),
               "rank": np.full(n_nodes, 1.0 / n_nodes, dtype=np.float32)},
    }


def bodies(n: int, seed: int = 71) -> dict:
    """Bodies as reffine's Nbody reads them (misc.h:63-210): id = row, mass
    in [1, 2), positions in [-1, 1)^3 (distinct with probability one),
    velocities in [-0.1, 0.1)^3."""
    u = lambda c: uniform(seed, c, n)
    return {"id": np.arange(n, dtype=np.int64), "m": 1.0 + u(1), "x": 2 * u(2) - 1,
            "y": 2 * u(3) - 1, "z": 2 * u(4) - 1, "vx": 0.2 * u(5) - 0.1,
            "vy": 0.2 * u(6) - 0.1, "vz": 0.2 * u(7) - 0.1}
