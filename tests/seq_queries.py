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
