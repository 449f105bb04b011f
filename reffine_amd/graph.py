"""reffine's PageRank step (benchmark/include/misc.h:212-291), as written.

`edges` and `rev_edges` are 2-dimensional tables whose column 0 (src, and
dst for the reverse) is run-end encoded: a node's out-edges (in-edges) are
one run. The function computes each node's contribution (its rank over its
out-degree, a fold of its edges run) as a named Op, builds an index over it
(`_buildidx`), and folds every node's in-edges run reading the contribution
of each edge's source from that index. The contribution table is computed
first, as a function of its own, and passed to the rest as one more table
(`split_materialized` in the lowering).
"""
from __future__ import annotations

from .reffiner import (_add, _buildidx, _cast, _div, _f32, _f32_t, _func, _get, _i64, _i64_t,
                       _in, _initval, _mul, _op, _red, _sym, _vec_t)

EDGES_TYPE = _vec_t(2, _i64_t, _i64_t, _i64_t, encodings=["runend", "flat", "flat"])
PR_TYPE = _vec_t(1, _i64_t, _f32_t)


def pagerank(n: int, alpha: float = 0.85):
    """misc.h:235-283 (PageRank::build_op): one step of
    new_pr[dst] = alpha * sum over in-edges of count * pr[src] / deg[src]
    + (1 - alpha) / n, in float32."""
    edges = _sym("edges", EDGES_TYPE)
    rev_edges = _sym("rev_edges", EDGES_TYPE)
    pr = _sym("pr", PR_TYPE)
    src = _sym("src", _i64_t)
    deg = _red(edges[src], lambda: _i64(0), lambda s, v: _add(s, _get(v, 1)))
    deg_sym = _sym("deg", deg)
    contrib = _div(_cast(_f32_t, _get(pr[src], 0)), _cast(_f32_t, deg_sym))
    contrib_sym = _sym("contrib", contrib)
    outdeg = _op([src], _in(src, edges) & _in(src, pr), [contrib_sym])
    outdeg_sym = _sym("outdeg", outdeg)
    outdeg2 = _buildidx(outdeg_sym)
    outdeg2_sym = _sym("outdeg2", outdeg2)
    dst = _sym("dst", _i64_t)

    def acc(s, v):
        return _add(s, _mul(_cast(_f32_t, _get(v, 1)), _get(outdeg2_sym[_get(v, 0)], 0)))

    dst_pr = _red(rev_edges[dst], lambda: _f32(0), acc)
    dst_pr_sym = _sym("dst_pr", dst_pr)
    new_pr_val = _add(_mul(_f32(alpha), dst_pr_sym), _f32((1 - alpha) / n))
    new_pr_val_sym = _sym("new_pr_val", new_pr_val)
    new_pr_op = _op([dst], _in(dst, rev_edges), [new_pr_val_sym])
    new_pr = _initval([outdeg2_sym], new_pr_op)
    new_pr_sym = _sym("new_pr", new_pr)
    fn = _func("pagerank", new_pr_sym, [edges, rev_edges, pr])
    fn.tbl[outdeg_sym] = outdeg
    fn.tbl[outdeg2_sym] = outdeg2
    fn.tbl[deg_sym] = deg
    fn.tbl[contrib_sym] = contrib
    fn.tbl[new_pr_sym] = new_pr
    fn.tbl[dst_pr_sym] = dst_pr
    fn.tbl[new_pr_val_sym] = new_pr_val
    return fn


def record_batches(tables: dict) -> dict:
    """pyarrow record batches: edges and rev_edges with column 0 run-end
    encoded (int32 run ends), pr flat."""
    import pyarrow as pa
    import pyarrow.compute as pc
    out = {}
    for name, cols in (("edges", ("src", "dst", "count")), ("rev_edges", ("dst", "src", "count"))):
        t = tables[name]
        arrays = [pc.run_end_encode(pa.array(t[cols[0]]), run_end_type=pa.int32())]
        arrays += [pa.array(t[c]) for c in cols[1:]]
        out[name] = pa.RecordBatch.from_arrays(arrays, names=list(cols))
    out["pr"] = pa.RecordBatch.from_arrays([pa.array(tables["pr"]["id"]),
                                            pa.array(tables["pr"]["rank"])], names=["id", "rank"])
    return out
