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

from .reffiner import (_add, _buildidx, _cast, _div, _f32, _f32_t, _f64_t, _func, _get, _i64, _i64_t,
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


BODIES_TYPE = _vec_t(1, _i64_t, _f64_t, _f64_t, _f64_t, _f64_t, _f64_t, _f64_t, _f64_t)


def nbody(n: int, g: float = 1.0, dt: float = 0.01):
    """misc.h:77-202 (Nbody::build_op), as written: the pairwise forces of
    n bodies as an Op over two ranges (id_a != id_b), then per body the sums
    of the pair table's rows [idx*n, (idx+1)*n) (the reference's windows:
    the pair table skips id_a == id_b, so a window runs into the next body's
    pairs, and the last one past the table's end, where the reference reads
    unchecked and this build folds only the rows that exist), then the
    updated positions and velocities. bodies: (id, m, x, y, z, vx, vy, vz)."""
    from .reffiner import _eq, _gte, _idx, _idx_t, _lt, _not, _readdata, _sqrt, _sub, _subvec, _f64
    bodies = _sym("bodies", BODIES_TYPE)
    id_a = _sym("id_a", _idx_t)
    id_b = _sym("id_b", _idx_t)
    cond = _not(_eq(id_a, id_b))
    cond_sym = _sym("cond", cond)
    dx = _sub(_readdata(bodies, id_a, 2), _readdata(bodies, id_b, 2))
    dx_sym = _sym("dx", dx)
    dy = _sub(_readdata(bodies, id_a, 3), _readdata(bodies, id_b, 3))
    dy_sym = _sym("dy", dy)
    dz = _sub(_readdata(bodies, id_a, 4), _readdata(bodies, id_b, 4))
    dz_sym = _sym("dz", dz)
    dist2 = _add(_add(_mul(dx_sym, dx_sym), _mul(dy_sym, dy_sym)), _mul(dz_sym, dz_sym))
    dist2_sym = _sym("dist2", dist2)
    dist = _sqrt(dist2_sym)
    dist_sym = _sym("dist", dist)
    f = _div(_mul(_f64(g), _mul(_readdata(bodies, id_a, 1), _readdata(bodies, id_b, 1))),
             dist2_sym)
    f_sym = _sym("f", f)
    fx = _div(_mul(f_sym, dx_sym), dist_sym)
    fx_sym = _sym("fx", fx)
    fy = _div(_mul(f_sym, dy_sym), dist_sym)
    fy_sym = _sym("fy", fy)
    fz = _div(_mul(f_sym, dz_sym), dist_sym)
    fz_sym = _sym("fz", fz)
    op = _op([id_a, id_b], cond_sym & _gte(id_a, _idx(0)) & _lt(id_a, _idx(n)) &
             _gte(id_b, _idx(0)) & _lt(id_b, _idx(n)), [fx_sym, fy_sym, fz_sym])
    op_sym = _sym("pairs", op)
    idx = _sym("idx", _idx_t)

    def window_sum(col):
        return _red(_subvec(op_sym, _mul(idx, _idx(n)), _mul(_add(idx, _idx(1)), _idx(n))),
                    lambda: _f64(0), lambda s, v: _add(s, _get(v, col)))

    sfx, sfy, sfz = window_sum(1), window_sum(2), window_sum(3)
    sfx_sym, sfy_sym, sfz_sym = _sym("sfx", sfx), _sym("sfy", sfy), _sym("sfz", sfz)
    op2 = _op([idx], _gte(idx, _idx(0)) & _lt(idx, _idx(n)), [sfx_sym, sfy_sym, sfz_sym])
    op2_2 = _initval([op_sym], op2)
    op2_sym = _sym("force", op2_2)
    id_ = _sym("ID", _idx_t)
    new_m_sym = _sym("M", _readdata(bodies, id_, 1))
    old_m = _readdata(bodies, id_, 1)
    vx = _add(_readdata(bodies, id_, 5),
              _mul(_div(_readdata(op2_sym, id_, 1), new_m_sym), _f64(dt)))
    vy = _add(_readdata(bodies, id_, 6),
              _mul(_div(_readdata(op2_sym, id_, 2), new_m_sym), _f64(dt)))
    vz = _add(_readdata(bodies, id_, 7),
              _mul(_div(_readdata(op2_sym, id_, 3), new_m_sym), _f64(dt)))
    vx_sym, vy_sym, vz_sym = _sym("VX", vx), _sym("VY", vy), _sym("VZ", vz)
    x = _add(_readdata(bodies, id_, 2), _mul(vx_sym, _f64(dt)))
    y = _add(_readdata(bodies, id_, 3), _mul(vy_sym, _f64(dt)))
    z = _add(_readdata(bodies, id_, 4), _mul(vz_sym, _f64(dt)))
    x_sym, y_sym, z_sym = _sym("X", x), _sym("Y", y), _sym("Z", z)
    op3 = _op([id_], _gte(id_, _idx(0)) & _lt(id_, _idx(n)),
              [new_m_sym, x_sym, y_sym, z_sym, vx_sym, vy_sym, vz_sym])
    op3_2 = _initval([op2_sym], op3)
    op3_sym = _sym("new_bodies", op3_2)
    fn = _func("nbodies", op3_sym, [bodies])
    for s_, e_ in ((new_m_sym, old_m), (vx_sym, vx), (vy_sym, vy), (vz_sym, vz), (x_sym, x),
                   (y_sym, y), (z_sym, z), (cond_sym, cond), (dx_sym, dx), (dy_sym, dy),
                   (dz_sym, dz), (dist2_sym, dist2), (dist_sym, dist), (f_sym, f), (fx_sym, fx),
                   (fy_sym, fy), (fz_sym, fz), (op_sym, op), (op2_sym, op2_2), (sfx_sym, sfx),
                   (sfy_sym, sfy), (sfz_sym, sfz), (op3_sym, op3_2)):
        fn.tbl[s_] = e_
    return fn
