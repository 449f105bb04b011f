"""Batched and served execution (the rf_batch_* and rf_serve C ABI).

A Batch collects (query, inputs) work units and runs them as one round:
one persistent work-queue launch per distinct query kernel, then per-unit
completion and selection placement. Each unit's result equals the
deterministic-mode run of the same query alone. `serve` runs a stream of
timed requests through two batch runners with continuous batching and
reports each request's completion time; `serve_plain` runs the same
requests one at a time through the plain per-query path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from .runtime import DeviceTable, Query, ReffineError, _ck, lib

RF_ERR_BUSY = 6


class RfRequest(ctypes.Structure):
    _fields_ = [("q", ctypes.c_void_p), ("args", ctypes.c_void_p), ("n_args", ctypes.c_int64),
                ("arrival_us", ctypes.c_double), ("output_addr", ctypes.c_void_p),
                ("sel_out", ctypes.c_void_p), ("sel_cap", ctypes.c_int64),
                ("out_rows", ctypes.c_int64), ("done_us", ctypes.c_double),
                ("status", ctypes.c_int32), ("pad", ctypes.c_int32)]


def _materialize(q: Query, raw: bytes, tbl: int | None):
    """Decodes an output; selection tables become host arrays (a batch's
    tables are views of its arena)."""
    if q.out_is_table:
        t = DeviceTable(tbl) if tbl else None
        if t is None:
            return None
        cols = [t.column(c) for c in range(t.ncols)]
        t.free()
        return cols
    val = q.decode(raw)

    def conv(v):
        if isinstance(v, tuple):
            return tuple(conv(x) for x in v)
        if isinstance(v, DeviceTable):
            cols = [v.column(c) for c in range(v.ncols)]
            v.owner = False  # the same view is released through out_table
            return cols
        return v
    out = conv(val)
    if tbl:
        DeviceTable(tbl).free()
    return out


class Batch:
    """One batch runner (rf_batch)."""

    def __init__(self, max_units: int = 256, arena_bytes: int = 1 << 30, use_graph: bool = False):
        h = ctypes.c_void_p()
        _ck(lib().rf_batch_create(max_units, arena_bytes, 1 if use_graph else 0, ctypes.byref(h)))
        self.handle = h
        self.units: list[tuple[Query, object]] = []

    def footprint(self, q: Query, *inputs) -> int:
        arr = q._args(inputs)
        b = ctypes.c_int64()
        _ck(lib().rf_batch_footprint(q.handle, arr, len(inputs), ctypes.byref(b)))
        return b.value

    def add(self, q: Query, *inputs) -> int | None:
        """Adds a unit; None when the round is full (units or arena)."""
        arr = q._args(inputs)
        u = ctypes.c_int32()
        st = lib().rf_batch_add(self.handle, q.handle, arr, len(inputs), ctypes.byref(u))
        if st == RF_ERR_BUSY:
            return None
        _ck(st)
        self.units.append((q, inputs))
        return u.value

    def launch(self):
        _ck(lib().rf_batch_launch(self.handle))

    def poll(self) -> bool:
        d = ctypes.c_int32()
        _ck(lib().rf_batch_poll(self.handle, ctypes.byref(d)))
        return bool(d.value)

    def wait(self):
        _ck(lib().rf_batch_wait(self.handle))

    def result(self, unit: int):
        q = self.units[unit][0]
        buf = ctypes.create_string_buffer(max(8, q.out_bytes))
        tbl = ctypes.c_void_p()
        _ck(lib().rf_batch_result(self.handle, unit, buf, ctypes.byref(tbl)))
        return _materialize(q, buf.raw[:q.out_bytes], tbl.value)

    def reset(self):
        _ck(lib().rf_batch_reset(self.handle))
        self.units.clear()

    def run(self, work: list[tuple[Query, tuple]]) -> list:
        """Runs every (query, inputs) unit in rounds; returns the outputs."""
        out, i = [], 0
        while i < len(work):
            start = i
            while i < len(work) and self.add(work[i][0], *work[i][1]) is not None:
                i += 1
            if i == start:
                raise ReffineError(RF_ERR_BUSY, "a unit does not fit an empty round")
            self.launch()
            self.wait()
            out += [self.result(u) for u in range(i - start)]
            self.reset()
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) and lib() is not None:
                lib().rf_batch_free(self.handle)
        except Exception:
            pass


@dataclass
class ServeResult:
    wall_us: float
    arrival_us: np.ndarray
    done_us: np.ndarray
    status: np.ndarray
    out_rows: np.ndarray
    outputs: list = field(default_factory=list)     # raw output bytes per request
    selections: list = field(default_factory=list)  # first selection column (bytes) per request

    @property
    def latency_us(self) -> np.ndarray:
        return self.done_us - self.arrival_us

    @property
    def qps(self) -> float:
        return len(self.done_us) / (self.wall_us * 1e-6)


def _requests(reqs, sel_cap_rows):
    """reqs: list of (Query, inputs tuple, arrival_us). sel_cap_rows: rows of
    the first selection column to copy back per request (an int for all, or
    a list; 0 copies none)."""
    n = len(reqs)
    arr = (RfRequest * max(n, 1))()
    keep = []
    for i, (q, inputs, t) in enumerate(reqs):
        a = q._args(inputs)
        ob = ctypes.create_string_buffer(max(8, q.out_bytes))
        cap = sel_cap_rows[i] if isinstance(sel_cap_rows, (list, tuple)) else sel_cap_rows
        sb = np.zeros(cap * 8, dtype=np.uint8) if cap > 0 else None
        keep += [a, ob, sb]
        r = arr[i]
        r.q = q.handle.value
        r.args = ctypes.cast(a, ctypes.c_void_p).value
        r.n_args = len(inputs)
        r.arrival_us = float(t)
        r.output_addr = ctypes.cast(ob, ctypes.c_void_p).value
        r.sel_out = sb.ctypes.data if sb is not None else None
        r.sel_cap = sb.nbytes if sb is not None else 0
    return arr, keep


def _collect(reqs, arr, keep, wall) -> ServeResult:
    n = len(reqs)
    res = ServeResult(wall_us=wall,
                      arrival_us=np.array([arr[i].arrival_us for i in range(n)]),
                      done_us=np.array([arr[i].done_us for i in range(n)]),
                      status=np.array([arr[i].status for i in range(n)], dtype=np.int32),
                      out_rows=np.array([arr[i].out_rows for i in range(n)], dtype=np.int64))
    for i, (q, _, _) in enumerate(reqs):
        ob, sb = keep[3 * i + 1], keep[3 * i + 2]
        res.outputs.append(ob.raw[:q.out_bytes])
        w = _sel_width(q)
        res.selections.append(sb[:arr[i].out_rows * w].tobytes() if sb is not None else None)
    return res


def _sel_width(q: Query) -> int:
    """Byte width of the first column of the query's selection output."""
    from .runtime import _layout

    def first_vec(t):
        if t.is_vector():
            return t
        if t.is_struct():
            for d in t.dtypes:
                v = first_vec(d)
                if v is not None:
                    return v
        return None
    v = first_vec(q.out_type)
    if v is None:
        return 0
    return _layout(v.dtypes[0])[0]


def serve(reqs, max_units: int = 256, arena_bytes: int = 1 << 30, use_graph: bool = True,
          sel_cap_rows=0, runners: int = 0) -> ServeResult:
    """Serves (query, inputs, arrival_us) requests with continuous batching
    over `runners` batch runners (2..4 rounds in flight; 0: the default, 3).
    The first selection column of each request is returned as bytes (at most
    sel_cap_rows rows of up to 8 bytes)."""
    arr, keep = _requests(reqs, sel_cap_rows)
    wall = ctypes.c_double()
    _ck(lib().rf_serve(arr, len(reqs), max_units, arena_bytes, 1 if use_graph else 0, runners,
                       ctypes.byref(wall)))
    return _collect(reqs, arr, keep, wall.value)


def serve_plain(reqs, sel_cap_rows=0) -> ServeResult:
    """The same requests, each run alone through the plain per-query path."""
    arr, keep = _requests(reqs, sel_cap_rows)
    wall = ctypes.c_double()
    _ck(lib().rf_serve_plain(arr, len(reqs), ctypes.byref(wall)))
    return _collect(reqs, arr, keep, wall.value)


def reduction_bytes(q: Query, raw: bytes) -> bytes:
    """The output bytes without embedded table pointers (they differ between
    runs); the reduction fields compare bit for bit."""
    t = q.out_type
    if not t.is_struct():
        return b"" if t.is_vector() else raw
    from .runtime import _layout
    keep = bytearray()

    def walk(tp, off):
        if tp.is_struct():
            for d in tp.dtypes:
                s, a = _layout(d)
                off = (off + a - 1) // a * a
                walk(d, off)
                off += s
            return
        s, _ = _layout(tp)
        if not tp.is_vector():
            keep.extend(raw[off:off + s])
    walk(t, 0)
    return bytes(keep)
