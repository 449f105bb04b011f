"""Python host side over the C ABI (include/reffine_hip.h).

Mirrors the reference's entry points for the scan path:

  * `compile_op(func, vectorize=False)` (include/reffine/utils/utils.h:47-54)
    returns a `Query`; calling it runs the lowered kernels on device tables,
    like calling the generated `void name(OutT*, ArrowTable*, ...)`
    (src/pass/cemitter.cpp:227-252);
  * `DeviceTable` stands where `ArrowTable2` / `load_arrow_file`
    (include/reffine/arrow/table.h:15-151, src/utils/utils.cpp:11-34) stand,
    with the buffers resident in HBM;
  * `DeviceTable.build_index()` / `locate()` are build_index /
    vector_locate (table.h:107-138, include/reffine/vinstr/vinstr.h:137-146).

The library must be built (`python -m reffine_amd.build`); nothing here
falls back to a CPU path.
"""
from __future__ import annotations

import ctypes
import struct
import os
from typing import Sequence

import numpy as np

from .reffiner import DataType, Func, _vec_t

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libreffine_hip.so")

RF_OK = 0
RF_COMPILE_DETERMINISTIC = 1


class ArrowSchema(ctypes.Structure):
    pass


ArrowSchema._fields_ = [
    ("format", ctypes.c_char_p), ("name", ctypes.c_char_p), ("metadata", ctypes.c_char_p),
    ("flags", ctypes.c_int64), ("n_children", ctypes.c_int64),
    ("children", ctypes.POINTER(ctypes.POINTER(ArrowSchema))),
    ("dictionary", ctypes.POINTER(ArrowSchema)), ("release", ctypes.c_void_p),
    ("private_data", ctypes.c_void_p)]


class ArrowArray(ctypes.Structure):
    pass


ArrowArray._fields_ = [
    ("length", ctypes.c_int64), ("null_count", ctypes.c_int64), ("offset", ctypes.c_int64),
    ("n_buffers", ctypes.c_int64), ("n_children", ctypes.c_int64),
    ("buffers", ctypes.POINTER(ctypes.c_void_p)),
    ("children", ctypes.POINTER(ctypes.POINTER(ArrowArray))),
    ("dictionary", ctypes.POINTER(ArrowArray)), ("release", ctypes.c_void_p),
    ("private_data", ctypes.c_void_p)]


class ArrowDeviceArray(ctypes.Structure):
    """Arrow C Device data interface: an ArrowArray and the device holding it."""
    _fields_ = [("array", ArrowArray), ("device_id", ctypes.c_int64),
                ("device_type", ctypes.c_int32), ("sync_event", ctypes.c_void_p),
                ("reserved", ctypes.c_int64 * 3)]


ARROW_DEVICE_CPU = 1
ARROW_DEVICE_ROCM = 10


def _release(struct_ptr, typ):
    """Calls an Arrow C struct's release callback (if it has one)."""
    rel = struct_ptr.contents.release
    if rel:
        ctypes.CFUNCTYPE(None, ctypes.POINTER(typ))(rel)(struct_ptr)


class _ArgU(ctypes.Union):
    _fields_ = [("table", ctypes.c_void_p), ("i64", ctypes.c_int64), ("f64", ctypes.c_double)]


class RfArg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("u", _ArgU)]


class ReffineError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


_lib = None

# (name, restype, argtypes) of every exported function.
_SIGS = [
    ("rf_init", ctypes.c_int, [ctypes.c_int]),
    ("rf_last_error", ctypes.c_char_p, []),
    ("rf_device_info", ctypes.c_int, [ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]),
    ("rf_sync", ctypes.c_int, []),
    ("rf_table_upload", ctypes.c_int, [ctypes.POINTER(ArrowSchema), ctypes.POINTER(ArrowArray),
                                       ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_table_import_device", ctypes.c_int, [ctypes.POINTER(ArrowSchema),
                                              ctypes.POINTER(ArrowDeviceArray), ctypes.c_int64,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_table_export_device", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ArrowSchema),
                                              ctypes.POINTER(ArrowDeviceArray)]),
    ("rf_table_wrap_device", ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_char_p),
                                            ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_table_shape", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]),
    ("rf_table_column", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p,
                                       ctypes.c_size_t]),
    ("rf_table_download_column", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                ctypes.c_size_t, ctypes.c_void_p,
                                                ctypes.c_size_t]),
    ("rf_table_build_index", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_table_locate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_void_p]),
    ("rf_table_free", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_compile", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_compile_tuned", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_compile_check", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p,
                                        ctypes.c_size_t]),
    ("rf_query_output_info", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int32)]),
    ("rf_query_describe", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("rf_query_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ("rf_query_clear_cache", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_query_run", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RfArg), ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_query_run_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RfArg),
                                            ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_void_p), ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]),
    ("rf_query_time", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RfArg), ctypes.c_int64,
                                     ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]),
    ("rf_query_free", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_calib_copy", ctypes.c_int, [ctypes.c_int64, ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_double)]),
    ("rf_batch_create", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_batch_footprint", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RfArg), ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_int64)]),
    ("rf_batch_add", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(RfArg),
                                    ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]),
    ("rf_batch_launch", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_batch_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ("rf_batch_wait", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_batch_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    ("rf_batch_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_batch_free", ctypes.c_int, [ctypes.c_void_p]),
    ("rf_serve", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("rf_serve_plain", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_double)]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib() -> ctypes.CDLL:
    """Loads the built library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m reffine_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ck(status: int):
    if status != RF_OK:
        raise ReffineError(status, lib().rf_last_error().decode(errors="replace"))


def init(device: int = 0) -> None:
    """CudaEngine::Get() counterpart (src/engine/cuda_engine.cpp:25-43)."""
    _ck(lib().rf_init(device))


def device_info() -> dict:
    n = ctypes.c_int32()
    tot = ctypes.c_int64()
    fr = ctypes.c_int64()
    _ck(lib().rf_device_info(ctypes.byref(n), ctypes.byref(tot), ctypes.byref(fr)))
    return {"num_cus": n.value, "total_bytes": tot.value, "free_bytes": fr.value}


# --------------------------------------------------------------------- tables

_NP_FORMAT = {np.dtype(np.int8): "c", np.dtype(np.uint8): "C", np.dtype(np.int16): "s",
              np.dtype(np.uint16): "S", np.dtype(np.int32): "i", np.dtype(np.uint32): "I",
              np.dtype(np.int64): "l", np.dtype(np.uint64): "L", np.dtype(np.float32): "f",
              np.dtype(np.float64): "g", np.dtype(np.bool_): "c"}
_FORMAT_NP = {"c": np.int8, "C": np.uint8, "s": np.int16, "S": np.uint16, "i": np.int32,
              "I": np.uint32, "l": np.int64, "L": np.uint64, "f": np.float32, "g": np.float64}
_FORMAT_TYPE = {"c": "i8", "C": "u8", "s": "i16", "S": "u16", "i": "i32", "I": "u32",
                "l": "i64", "L": "u64", "f": "f32", "g": "f64"}


class _HostBatch:
    """Arrow C-data struct array over numpy columns (kept alive while used)."""

    def __init__(self, columns: Sequence[np.ndarray], validity, names):
        self.keep = []
        n = len(columns)
        self.schema = ArrowSchema()
        self.array = ArrowArray()
        child_s = (ctypes.POINTER(ArrowSchema) * n)()
        child_a = (ctypes.POINTER(ArrowArray) * n)()
        length = len(columns[0]) if n else 0
        for i, col in enumerate(columns):
            col = np.ascontiguousarray(col)
            if col.dtype == np.bool_:
                col = col.astype(np.int8)
            if col.dtype not in _NP_FORMAT:
                raise TypeError(f"column {i}: unsupported dtype {col.dtype}")
            if len(col) != length:
                raise ValueError("columns differ in length")
            s = ArrowSchema()
            s.format = _NP_FORMAT[col.dtype].encode()
            s.name = (names[i] if names else f"c{i}").encode()
            a = ArrowArray()
            a.length = len(col)
            a.offset = 0
            a.n_buffers = 2
            bufs = (ctypes.c_void_p * 2)()
            vb = None
            if validity is not None and validity[i] is not None:
                v = np.asarray(validity[i], dtype=bool)
                vb = np.packbits(v, bitorder="little")
                a.null_count = int(len(v) - v.sum())
                bufs[0] = vb.ctypes.data
            else:
                a.null_count = 0
                bufs[0] = None
            bufs[1] = col.ctypes.data
            a.buffers = bufs
            self.keep += [s, a, bufs, col, vb]
            child_s[i] = ctypes.pointer(s)
            child_a[i] = ctypes.pointer(a)
        self.schema.format = b"+s"
        self.schema.n_children = n
        self.schema.children = child_s
        self.array.length = length
        self.array.n_children = n
        self.array.n_buffers = 1
        top = (ctypes.c_void_p * 1)()
        self.array.buffers = top
        self.array.children = child_a
        self.keep += [child_s, child_a, top]


class DeviceTable:
    """A record batch resident in HBM (one device allocation per table)."""

    def __init__(self, handle: int, owner: bool = True):
        self.handle = ctypes.c_void_p(handle)
        self.owner = owner

    def __getattr__(self, name):
        # The shape is read from the library on first use (a query result
        # that is only passed on to another query never needs it).
        if name in ("nrows", "ncols", "formats", "lengths"):
            self._describe()
            return self.__dict__[name]
        raise AttributeError(name)

    def _describe(self):
        nr = ctypes.c_int64()
        nc = ctypes.c_int64()
        _ck(lib().rf_table_shape(self.handle, ctypes.byref(nr), ctypes.byref(nc)))
        self.nrows, self.ncols = nr.value, nc.value
        self.formats = []
        self.lengths = []
        for c in range(self.ncols):
            fmt = ctypes.create_string_buffer(16)
            ln = ctypes.c_int64()
            _ck(lib().rf_table_column(self.handle, c, None, None, ctypes.byref(ln), fmt, 16))
            self.formats.append(fmt.value.decode())
            self.lengths.append(ln.value)

    @classmethod
    def from_columns(cls, columns: Sequence[np.ndarray], dim: int = 1,
                     validity: Sequence[np.ndarray | None] | None = None,
                     names: Sequence[str] | None = None) -> "DeviceTable":
        hb = _HostBatch(columns, validity, names)
        out = ctypes.c_void_p()
        _ck(lib().rf_table_upload(ctypes.byref(hb.schema), ctypes.byref(hb.array), dim,
                                  ctypes.byref(out)))
        t = cls(out.value)
        t.dim = dim
        return t

    @classmethod
    def from_arrow(cls, batch, dim: int = 1) -> "DeviceTable":
        """Uploads a pyarrow RecordBatch through the Arrow C data interface."""
        s = ArrowSchema()
        a = ArrowArray()
        batch._export_to_c(ctypes.addressof(a), ctypes.addressof(s))
        out = ctypes.c_void_p()
        try:
            _ck(lib().rf_table_upload(ctypes.byref(s), ctypes.byref(a), dim, ctypes.byref(out)))
        finally:
            if a.release:
                ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowArray))(a.release)(ctypes.byref(a))
            if s.release:
                ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowSchema))(s.release)(ctypes.byref(s))
        t = cls(out.value)
        t.dim = dim
        return t

    @classmethod
    def from_arrow_device(cls, schema: "ArrowSchema", array: "ArrowDeviceArray",
                          dim: int = 1) -> "DeviceTable":
        """Imports an Arrow C Device array (moved in: its release callback
        runs here or when the table is freed; the schema stays the caller's).
        ROCm memory is borrowed, host memory uploaded."""
        out = ctypes.c_void_p()
        _ck(lib().rf_table_import_device(ctypes.byref(schema), ctypes.byref(array), dim,
                                         ctypes.byref(out)))
        t = cls(out.value)
        t.dim = dim
        return t

    @classmethod
    def from_arrow_c_device(cls, batch, dim: int = 1) -> "DeviceTable":
        """Uploads a pyarrow RecordBatch through the Arrow C Device data
        interface (pyarrow exports host batches as device type CPU)."""
        s = ArrowSchema()
        a = ArrowDeviceArray()
        batch._export_to_c_device(ctypes.addressof(a), ctypes.addressof(s))
        try:
            return cls.from_arrow_device(s, a, dim)
        finally:
            _release(ctypes.pointer(s), ArrowSchema)

    def export_device(self) -> "DeviceExport":
        """This table as an Arrow C Device array (ROCm, zero copy)."""
        e = DeviceExport()
        _ck(lib().rf_table_export_device(self.handle, ctypes.byref(e.schema),
                                         ctypes.byref(e.array)))
        return e

    @property
    def dtype(self) -> DataType:
        """The reffine vector type (ArrowTable2::get_data_type, table.h:64-80)."""
        dim = getattr(self, "dim", 1)
        return _vec_t(dim, *[DataType(_FORMAT_TYPE[f]) for f in self.formats])

    def column(self, c: int) -> np.ndarray:
        out = np.empty(self.lengths[c], dtype=_FORMAT_NP[self.formats[c]])
        _ck(lib().rf_table_download_column(self.handle, c, out.ctypes.data, out.nbytes, None, 0))
        return out

    def validity(self, c: int) -> np.ndarray:
        n = self.lengths[c]
        out = np.empty(self.lengths[c], dtype=_FORMAT_NP[self.formats[c]])
        vb = np.empty((n + 7) // 8, dtype=np.uint8)
        _ck(lib().rf_table_download_column(self.handle, c, out.ctypes.data, out.nbytes,
                                           vb.ctypes.data, vb.nbytes))
        return np.unpackbits(vb, bitorder="little")[:n].astype(bool)

    def device_ptr(self, c: int) -> int:
        p = ctypes.c_void_p()
        _ck(lib().rf_table_column(self.handle, c, ctypes.byref(p), None, None, None, 0))
        return p.value or 0

    def view(self, dim: int) -> "DeviceTable":
        """The same device columns as a table of another iterator dimension
        (e.g. dim 2 for a `_subvec` fold and dim 1 for an `_op` over the
        rows). Borrows the buffers: this table must outlive the view."""
        n = self.ncols
        fmts = (ctypes.c_char_p * n)(*[f.encode() for f in self.formats])
        vals = (ctypes.c_void_p * n)()
        valid = (ctypes.c_void_p * n)()
        for c in range(n):
            v, b = ctypes.c_void_p(), ctypes.c_void_p()
            _ck(lib().rf_table_column(self.handle, c, ctypes.byref(v), ctypes.byref(b), None, None, 0))
            vals[c], valid[c] = v.value, b.value
        out = ctypes.c_void_p()
        _ck(lib().rf_table_wrap_device(dim, n, self.nrows, fmts, vals, valid, ctypes.byref(out)))
        t = DeviceTable(out.value)
        t.dim = dim
        t._base = self
        return t

    def build_index(self) -> "DeviceTable":
        _ck(lib().rf_table_build_index(self.handle))
        return self

    def free(self):
        if self.handle and self.owner:
            _ck(lib().rf_table_free(self.handle))
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.handle and self.owner and _lib is not None:
                _lib.rf_table_free(self.handle)
        except Exception:
            pass


# -------------------------------------------------------------------- queries

def _layout(t: DataType):
    """C layout (size, align) of an output type: natural alignment, fields in
    declaration order (ScalarPass, src/pass/scalarpass.cpp:103-227)."""
    if t.is_struct():
        off, al = 0, 1
        for d in t.dtypes:
            s, a = _layout(d)
            off = (off + a - 1) // a * a + s
            al = max(al, a)
        return (off + al - 1) // al * al, al
    if t.is_vector():
        return 8, 8
    w = {"b": 1, "i8": 1, "u8": 1, "i16": 2, "u16": 2, "i32": 4, "u32": 4, "f32": 4,
         "i64": 8, "u64": 8, "f64": 8, "x": 8}[t.btype]
    return w, w


_NP_OF = {"b": np.bool_, "i8": np.int8, "u8": np.uint8, "i16": np.int16, "u16": np.uint16,
          "i32": np.int32, "u32": np.uint32, "i64": np.int64, "u64": np.uint64,
          "f32": np.float32, "f64": np.float64, "x": np.int64}


_SC = {"b": "?", "i8": "b", "u8": "B", "i16": "h", "u16": "H", "i32": "i", "u32": "I",
       "f32": "f", "i64": "q", "u64": "Q", "f64": "d", "x": "q"}


class _Decoder:
    """Reads an output struct with one struct.unpack_from (fields at their C
    offsets, padding skipped) and nests the values as the type does."""

    def __init__(self, t: DataType):
        fmt, leaves, self.vec = ["<"], [], []
        pos = 0

        def walk(tp, off):
            nonlocal pos
            if tp.is_struct():
                out = []
                for d in tp.dtypes:
                    sz, al = _layout(d)
                    off = (off + al - 1) // al * al
                    out.append(walk(d, off))
                    off += sz
                return tuple(out)
            if off > pos:
                fmt.append(f"{off - pos}x")
            sz, _ = _layout(tp)
            fmt.append("Q" if tp.is_vector() else _SC[tp.btype])
            pos = off + sz
            leaves.append(tp.is_vector())
            return len(leaves) - 1

        self.tpl = walk(t, 0)
        self.vec = leaves
        self.st = struct.Struct("".join(fmt))

    def __call__(self, buf: bytes):
        vals = self.st.unpack_from(buf)

        def build(x):
            if isinstance(x, tuple):
                return tuple(build(y) for y in x)
            v = vals[x]
            if self.vec[x]:
                return DeviceTable(v) if v else None
            return v
        return build(self.tpl)


def _decode(buf: bytes, t: DataType, off: int = 0):
    """Decodes an output struct at offset `off` of `buf`."""
    return _Decoder(t)(buf[off:])


class DeviceExport:
    """An exported (schema, ArrowDeviceArray) pair; releases both when
    collected unless moved (import_as_table)."""

    def __init__(self):
        self.schema = ArrowSchema()
        self.array = ArrowDeviceArray()

    def import_as_table(self, dim: int = 1) -> "DeviceTable":
        """Moves the array into a new device table that borrows its buffers."""
        return DeviceTable.from_arrow_device(self.schema, self.array, dim)

    def __del__(self):
        try:
            if self.array.array.release:
                ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowArray))(self.array.array.release)(
                    ctypes.byref(self.array.array))
            if self.schema.release:
                ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowSchema))(self.schema.release)(
                    ctypes.byref(self.schema))
        except Exception:
            pass


class Query:
    """A compiled query: the GPU counterpart of compile_op's function pointer."""

    def __init__(self, func: Func, vectorize: bool = False, deterministic: bool = False,
                 tuning: tuple | None = None, index_scale: int | None = None):
        self.func = func
        self.vectorize = vectorize
        self.deterministic = deterministic
        self.index_scale = index_scale
        self.text = func.ir()
        h = ctypes.c_void_p()
        flags = RF_COMPILE_DETERMINISTIC if deterministic else 0
        if index_scale:
            flags |= int(index_scale) << 8  # RF_COMPILE_INDEX_SCALE
        if tuning is None:
            _ck(lib().rf_compile(self.text.encode(), flags, ctypes.byref(h)))
        else:
            tv = (ctypes.c_int32 * 6)(*(list(tuning) + [0] * (6 - len(tuning))))
            _ck(lib().rf_compile_tuned(self.text.encode(), flags, tv, ctypes.byref(h)))
        self.handle = h
        nb = ctypes.c_int64()
        it = ctypes.c_int32()
        _ck(lib().rf_query_output_info(h, ctypes.byref(nb), ctypes.byref(it)))
        self.out_bytes, self.out_is_table = nb.value, bool(it.value)
        self.out_type = func.output.type

    def _args(self, inputs):
        last = self.__dict__.get("_last_args")
        if last is not None and len(last[0]) == len(inputs) and all(
                a is b or (not isinstance(a, DeviceTable) and type(a) is type(b) and a == b)
                for a, b in zip(last[0], inputs)):
            return last[1]  # the same inputs as the previous call (held, so ids stay unique)
        if len(inputs) != len(self.func.inputs):
            raise TypeError(f"{self.func.name} takes {len(self.func.inputs)} inputs")
        arr = (RfArg * max(1, len(inputs)))()
        for i, (x, sym) in enumerate(zip(inputs, self.func.inputs)):
            if sym.type.is_vector():
                if not isinstance(x, DeviceTable):
                    raise TypeError(f"input {i} must be a DeviceTable")
                arr[i].kind = 0
                arr[i].u.table = x.handle.value
            elif sym.type.is_float():
                arr[i].kind = 2
                arr[i].u.f64 = float(x)
            else:
                arr[i].kind = 1
                arr[i].u.i64 = int(x)
        self._last_args = (tuple(inputs), arr)
        return arr

    def run_raw(self, *inputs) -> tuple[bytes, DeviceTable | None]:
        arr = self._args(inputs)
        buf = ctypes.create_string_buffer(max(8, self.out_bytes))
        tbl = ctypes.c_void_p()
        # A struct output carries its Op tables inside the struct (decoded
        # by Query.decode); only a vector output comes back through out_table.
        _ck(lib().rf_query_run(self.handle, arr, len(inputs), buf,
                               ctypes.byref(tbl) if self.out_is_table else None))
        return buf.raw[:self.out_bytes], (DeviceTable(tbl.value) if tbl.value else None)

    def __call__(self, *inputs):
        raw, tbl = self.run_raw(*inputs)
        if self.out_is_table:
            return tbl
        return self.decode(raw)

    def decode(self, raw: bytes):
        """The function output from the bytes of its output struct."""
        dec = getattr(self, "_decoder", None)
        if dec is None:
            dec = self._decoder = _Decoder(self.out_type)
        return dec(raw)

    def time(self, *inputs, iters: int = 20) -> float:
        """Mean device ms of a run (all of the query's kernels)."""
        return self.time_kernels(*inputs, iters=iters)[0]

    def time_kernels(self, *inputs, iters: int = 20) -> tuple[float, float]:
        """(mean device ms of a run, mean ms of the kernel that streams the
        input table alone: the scan kernel, or a run stage's run fold), from
        HIP events on the library stream."""
        arr = self._args(inputs)
        ms, scan = ctypes.c_double(), ctypes.c_double()
        _ck(lib().rf_query_time(self.handle, arr, len(inputs), iters, ctypes.byref(ms),
                                ctypes.byref(scan)))
        return ms.value, scan.value

    def tuning(self) -> tuple[int, int, int, int, int, int]:
        """(BLOCK, ITEMS, SUB, WAVES_PER_CU, PIPELINE, LSUB) of the scan
        kernel; with deterministic=True BLOCK, ITEMS and SUB fix the
        reduction order."""
        t = (ctypes.c_int32 * 6)()
        _ck(lib().rf_query_tuning(self.handle, t))
        return tuple(t)

    def clear_cache(self):
        """Drops a group join's resident build sides (the next run builds
        again)."""
        _ck(lib().rf_query_clear_cache(self.handle))

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1 << 20)
        _ck(lib().rf_query_describe(self.handle, buf, len(buf)))
        return buf.value.decode()

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.rf_query_free(self.handle)
        except Exception:
            pass


def compile_op(func: Func, vectorize: bool = False, deterministic: bool = False,
               tuning: tuple | None = None, index_scale: int | None = None) -> Query:
    """compile_op<T>(func, vectorize) (include/reffine/utils/utils.h:47-54).
    `tuning` = (BLOCK, ITEMS, SUB[, WAVES_PER_CU[, PIPELINE[, LSUB]]]) overrides the
    default launch shape; `index_scale` (2..15) the slots per key of a group
    join's hash index over sparse build keys."""
    return Query(func, vectorize, deterministic, tuning, index_scale)


def calib_copy(nbytes: int, iters: int = 20) -> float:
    """Mean ms of a plain device copy of nbytes (read nbytes + write nbytes)."""
    ms = ctypes.c_double()
    _ck(lib().rf_calib_copy(nbytes, iters, ctypes.byref(ms)))
    return ms.value


def compile_check(func: Func, deterministic: bool = False, nullable: bool = False) -> str:
    """Lowers and hiprtc-compiles a Func for gfx950 without touching a GPU."""
    buf = ctypes.create_string_buffer(1 << 20)
    flags = (RF_COMPILE_DETERMINISTIC if deterministic else 0) | (2 if nullable else 0)
    _ck(lib().rf_compile_check(func.ir().encode(), flags, buf, len(buf)))
    return buf.value.decode()


def load_arrow_file(path: str, dim: int = 1) -> DeviceTable:
    """Reads an Arrow IPC file (file or stream format) and uploads it as one
    device table -- the counterpart of the reference's load_arrow_file
    (src/utils/utils.cpp:11-34), which reads the same files into a host
    ArrowTable2. Record batches of the file are combined into one."""
    import pyarrow as pa
    import pyarrow.ipc as ipc
    try:
        with ipc.open_file(path) as r:
            table = r.read_all()
    except pa.ArrowInvalid:
        with ipc.open_stream(path) as r:
            table = r.read_all()
    batches = table.combine_chunks().to_batches()
    if not batches:
        batch = pa.RecordBatch.from_arrays([pa.array([], type=f.type) for f in table.schema],
                                           schema=table.schema)
    else:
        batch = batches[0]
    return DeviceTable.from_arrow(batch, dim)
