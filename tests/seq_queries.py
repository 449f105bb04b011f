"""Folds the scan kernels cannot re-associate (each field's next value is
not `pred ? s (+) x : s` of a row value): the library runs them as one
sequential fold in row order. Shared by the CPU oracle checks and the GPU
parity tests."""
import numpy as np

from reffine_amd.reffiner import (_add, _f64, _f64_t, _func, _get, _gt, _i64, _i64_t, _idx,
                                  _len, _max, _mul, _new, _red, _sel, _subvec, _sym, _vec_t)

SEQ_T = _vec_t(2, _i64_t, _i64_t, _f64_t)  # rowid, k, v


def _fold(name, init, acc):
    T = _sym("T", SEQ_T)
    red = _red(_subvec(T, _idx(0), _len(T, 1)), init, acc)
    rs = _sym("r", red)
    fn = _func(name, rs, [T])
    fn.tbl[rs] = red
    return fn


def decay():
    """s = s * 0.5 + v: an exponentially decayed sum."""
    return _fold("decay", lambda: _f64(0.25), lambda s, r: _add(_mul(s, _f64(0.5)), _get(r, 1)))


def poly_hash():
    """s = v > 0.5 ? s * 3 + k : s (int64, wrapping)."""
    return _fold("poly", lambda: _i64(7),
                 lambda s, r: _sel(_gt(_get(r, 1), _f64(0.5)),
                                   _add(_mul(s, _i64(3)), _get(r, 0)), s))


def max_state_second():
    """s = std::max(v, s): with the state second, a NaN row replaces the
    state and the next row replaces the NaN."""
    return _fold("maxv", lambda: _f64(-1.0), lambda s, r: _max(_get(r, 1), s))


def running_pair():
    """{a, b} = {a + v, b + a}: the second field reads the first."""
    return _fold("pair", lambda: _new([_f64(0), _f64(0)]),
                 lambda s, r: _new([_add(_get(s, 0), _get(r, 1)),
                                    _add(_get(s, 1), _get(s, 0))]))


def loop(name, k, v, valid=None):
    """The same folds as plain row loops in Python (IEEE doubles; int64
    wrapping), rows in order, a row whose column 1 is null skipped."""
    n = len(v)
    ok = np.ones(n, bool) if valid is None else valid
    if name == "decay":
        s = 0.25
        for i in range(n):
            if ok[i]:
                s = s * 0.5 + float(v[i])
        return (s,)
    if name == "poly":
        s = np.int64(7)
        with np.errstate(over="ignore"):
            for i in range(n):
                if ok[i] and float(v[i]) > 0.5:
                    s = np.int64(s * np.int64(3) + np.int64(k[i]))
        return (int(s),)
    if name == "maxv":
        s = -1.0
        for i in range(n):
            if ok[i]:
                x = float(v[i])
                s = s if x < s else x  # std::max(x, s)
        return (s,)
    if name == "pair":
        a = b = 0.0
        for i in range(n):
            if ok[i]:
                a, b = a + float(v[i]), b + a
        return (a, b)
    raise KeyError(name)


BUILDERS = {"decay": decay, "poly": poly_hash, "maxv": max_state_second, "pair": running_pair}


def _red_sum(T, col):
    return _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _f64(0.0),
                lambda s, r: _add(s, _get(r, col)))


def sum_and_count_of_two_tables():
    """{SUM(T.v), COUNT(U)}: two reductions over different tables (each its
    own query: they cannot share a scan)."""
    T = _sym("T", SEQ_T)
    U = _sym("U", _vec_t(2, _i64_t, _f64_t))
    a = _red_sum(T, 1)
    b = _red(_subvec(U, _idx(0), _len(U, 1)), lambda: _i64(0),
             lambda s, r: _add(s, _i64(1)))
    sa, sb = _sym("sa", a), _sym("sb", b)
    out = _new([sa, sb])
    so = _sym("out", out)
    fn = _func("two_tables", so, [T, U])
    fn.tbl[sa] = a
    fn.tbl[sb] = b
    fn.tbl[so] = out
    return fn


def decay_and_sum():
    """{decayed sum (sequential), SUM(v)} over one table: a sequential fold
    beside an associative one."""
    T = _sym("T", SEQ_T)
    d = _red(_subvec(T, _idx(0), _len(T, 1)), lambda: _f64(0.25),
             lambda s, r: _add(_mul(s, _f64(0.5)), _get(r, 1)))
    a = _red_sum(T, 1)
    sd, sa = _sym("sd", d), _sym("sa", a)
    out = _new([sd, sa])
    so = _sym("out", out)
    fn = _func("decay_and_sum", so, [T])
    fn.tbl[sd] = d
    fn.tbl[sa] = a
    fn.tbl[so] = out
    return fn


def selection_and_decay():
    """{rows with v > 0.9 (an Op), decayed sum (sequential)} over one table."""
    from reffine_amd.reffiner import _op, _in
    T = _sym("T1", _vec_t(1, _i64_t, _i64_t, _f64_t))
    t = _sym("t", _i64_t)
    op = _op([t], _in(t, T) & _gt(T[t][1], _f64(0.9)), [T[t][1]])
    t2 = _sym("t2", _i64_t)
    rows = _op([t2], _in(t2, T), [T[t2][1]])
    d = _red(rows, lambda: _f64(0.25), lambda s, r: _add(_mul(s, _f64(0.5)), _get(r, 1)))
    so_, sd = _sym("sel", op), _sym("sd", d)
    out = _new([so_, sd])
    so = _sym("out", out)
    fn = _func("selection_and_decay", so, [T])
    fn.tbl[so_] = op
    fn.tbl[sd] = d
    fn.tbl[so] = out
    return fn
