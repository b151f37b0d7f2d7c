"""ctypes binding of the ASC::sparse_hip C ABI (include/asc_sparse_hip.h).

This is the binding a Python caller would write against the C ABI: plain
structs, raw device pointers, status codes mapped to exceptions.  It loads the
in-tree ``lib/libasc_sparse_hip.so`` and fails loudly when it is missing --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
# ASC_SPARSE_HIP_LIBRARY names another build of the C ABI library (the host
# sanitizer build of tests/test_sanitized_abi.py); the default is the in-tree one.
LIB_PATH = os.environ.get("ASC_SPARSE_HIP_LIBRARY") or os.path.join(LIB_DIR, "libasc_sparse_hip.so")
GEN_PATH = os.path.join(LIB_DIR, "libasc_sparse_gen.so")
PROBE_PATH = os.path.join(LIB_DIR, "libasc_sparse_probe.so")

# ---- enums (values of include/asc_sparse_hip.h) ---------------------------
OK, INVALID_ARGUMENT, SHAPE, INDEX, OVERFLOW, INVALID_STATE, ALLOCATION = 0, 1, 2, 3, 4, 5, 6
MEMORY_ACCESS, MEMORY_TRANSFER, UNSUPPORTED, UNAVAILABLE, PROVIDER = 7, 8, 9, 10, 11
NUMERICAL, INTERNAL = 12, 18

F32, F64, C64, C128, BF16, F16 = 0, 1, 2, 3, 4, 5
COO, CSR, CSC = 0, 1, 2
HOST, PINNED_HOST, DEVICE, MANAGED = 0, 1, 2, 3
OP_N, OP_T, OP_H = 0, 1, 2
COLUMN_MAJOR, ROW_MAJOR = 0, 1
DETERMINISTIC, BACKEND_DEFAULT = 0, 1

ALG_AUTO, ALG_ORDERED, ALG_MERGE = 0, 1, 2
ALG_VECTOR = {1: 3, 2: 4, 4: 5, 8: 6, 16: 7, 32: 8, 64: 9}
ALG_PANEL = 10
ALG_NAMES = {2: "merge", 3: "vec1", 4: "vec2", 5: "vec4", 6: "vec8", 7: "vec16",
             8: "vec32", 9: "vec64", 1: "ordered", 10: "panel"}
TUNABLE_ALGS = (2, 3, 4, 5, 6, 7, 8, 9)
ACC_PLAIN, ACC_TWOSUM = 0, 1
INJECT_ALLOCATION, INJECT_EVENT, INJECT_STREAM_STALL = 0, 1, 2

ELEM_BYTES = {F32: 4, F64: 8, C64: 8, C128: 16, BF16: 2, F16: 2}
NP_STORAGE = {F32: np.float32, F64: np.float64, BF16: np.uint16, F16: np.float16}


class Status(C.Structure):
    _fields_ = [("code", C.c_int32), ("provider", C.c_int32), ("native", C.c_int64)]


class SparseDesc(C.Structure):
    _fields_ = [("outer_offsets", C.c_void_p), ("inner_indices", C.c_void_p),
                ("values", C.c_void_p), ("rows", C.c_int64), ("cols", C.c_int64),
                ("nnz", C.c_int64), ("format", C.c_uint8), ("element_kind", C.c_uint8),
                ("index_bytes", C.c_uint8), ("offset_bytes", C.c_uint8),
                ("memory_space", C.c_uint8), ("trusted", C.c_uint8),
                ("reserved", C.c_uint8 * 2)]


class VectorDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("extent", C.c_int64), ("stride", C.c_int64),
                ("element_kind", C.c_uint8), ("memory_space", C.c_uint8),
                ("writable", C.c_uint8), ("reserved", C.c_uint8 * 5)]


class MatrixDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("rows", C.c_int64), ("cols", C.c_int64),
                ("leading_dimension", C.c_int64), ("layout", C.c_uint8),
                ("element_kind", C.c_uint8), ("memory_space", C.c_uint8),
                ("writable", C.c_uint8), ("reserved", C.c_uint8 * 4)]


class Scalar(C.Structure):
    _fields_ = [("real", C.c_double), ("imaginary", C.c_double)]


class SpmvOptions(C.Structure):
    _fields_ = [("algorithm", C.c_uint8), ("accumulate", C.c_uint8),
                ("reserved", C.c_uint8 * 6)]


class BatchEntry(C.Structure):
    _fields_ = [("matrix", SparseDesc), ("x", VectorDesc), ("y", VectorDesc),
                ("alpha", C.c_double), ("beta", C.c_double)]


class BatchClasses(C.Structure):
    _fields_ = [("lanes", C.c_uint8 * 8)]


class BatchPlan(C.Structure):
    _fields_ = [("matrix_table", C.c_void_p), ("item_table", C.c_void_p),
                ("matrices", C.c_int64), ("items", C.c_int64), ("grid", C.c_int64),
                ("device", C.c_int32), ("element_kind", C.c_uint8), ("index_bytes", C.c_uint8),
                ("reserved", C.c_uint8 * 2)]


class AscHipError(RuntimeError):
    def __init__(self, code: int, message: str, native: int = 0):
        super().__init__(f"asc_sparse_hip error {code}: {message}")
        self.code = code
        self.native = native
        self.message = message


P = C.POINTER
_VP = C.c_void_p
_SIGS = {
    "asc_sparse_hip_context_create": [C.c_int32, _VP, C.c_uint8, P(_VP)],
    "asc_sparse_hip_context_destroy": [_VP],
    "asc_sparse_hip_context_stream": [_VP, P(_VP)],
    "asc_sparse_hip_context_synchronize": [_VP],
    "asc_sparse_hip_allocate": [_VP, C.c_size_t, C.c_uint8, P(_VP)],
    "asc_sparse_hip_deallocate": [_VP, _VP, C.c_uint8],
    "asc_sparse_hip_copy": [_VP, _VP, _VP, C.c_size_t],
    "asc_sparse_hip_fill": [_VP, _VP, C.c_uint8, C.c_size_t],
    "asc_sparse_hip_event_record": [_VP, P(_VP)],
    "asc_sparse_hip_event_wait": [_VP],
    "asc_sparse_hip_event_elapsed_ms": [_VP, _VP, P(C.c_float)],
    "asc_sparse_hip_event_destroy": [_VP],
    "asc_sparse_hip_clone_compressed": [_VP, P(SparseDesc), C.c_uint8, _VP, _VP, _VP, P(SparseDesc)],
    "asc_sparse_hip_clone_coordinate": [_VP, P(SparseDesc), _VP, _VP, P(SparseDesc)],
    "asc_sparse_hip_rebind_values": [P(SparseDesc), _VP, P(SparseDesc)],
    "asc_sparse_hip_narrow_indices": [_VP, P(SparseDesc), _VP, P(SparseDesc)],
    "asc_sparse_hip_convert_values": [_VP, P(SparseDesc), C.c_uint8, _VP, P(SparseDesc)],
    "asc_sparse_hip_convert_values_scaled": [_VP, P(SparseDesc), C.c_uint8, C.c_double, _VP, P(SparseDesc)],
    "asc_sparse_hip_csr_spmv_workspace_size": [_VP, P(SparseDesc), P(VectorDesc), P(VectorDesc), P(SpmvOptions), P(C.c_size_t)],
    "asc_sparse_hip_csr_spmv": [_VP, C.c_double, P(SparseDesc), P(VectorDesc), C.c_double, P(VectorDesc), _VP, C.c_size_t, P(SpmvOptions)],
    "asc_sparse_hip_csr_spmv_noalloc": [_VP, C.c_double, P(SparseDesc), P(VectorDesc), C.c_double, P(VectorDesc), P(SpmvOptions)],
    "asc_sparse_hip_standard_spmv_workspace_size": [_VP, C.c_uint8, P(SparseDesc), P(VectorDesc), P(VectorDesc), P(SpmvOptions), P(C.c_size_t)],
    "asc_sparse_hip_standard_spmv": [_VP, C.c_uint8, Scalar, P(SparseDesc), P(VectorDesc), P(VectorDesc), _VP, C.c_size_t, P(SpmvOptions)],
    "asc_sparse_hip_standard_spmm_workspace_size": [_VP, C.c_uint8, P(SparseDesc), P(MatrixDesc), P(MatrixDesc), P(SpmvOptions), P(C.c_size_t)],
    "asc_sparse_hip_standard_spmm": [_VP, C.c_uint8, Scalar, P(SparseDesc), P(MatrixDesc), P(MatrixDesc), _VP, C.c_size_t, P(SpmvOptions)],
    "asc_sparse_hip_convert_coordinate_workspace_size": [_VP, P(SparseDesc), C.c_uint8, C.c_uint8, P(C.c_size_t)],
    "asc_sparse_hip_convert_coordinate": [_VP, P(SparseDesc), C.c_uint8, C.c_uint8, _VP, _VP, _VP, P(SparseDesc), _VP, C.c_size_t],
    "asc_sparse_hip_convert_compressed_workspace_size": [_VP, P(SparseDesc), P(C.c_size_t)],
    "asc_sparse_hip_convert_compressed": [_VP, P(SparseDesc), _VP, _VP, _VP, P(SparseDesc), _VP, C.c_size_t],
    "asc_sparse_hip_test_inject": [C.c_uint32, C.c_int32],
    "asc_sparse_hip_batch_plan_size": [_VP, C.c_int64, P(BatchEntry), P(C.c_size_t)],
    "asc_sparse_hip_batch_plan_build": [_VP, C.c_int64, P(BatchEntry), P(BatchClasses), _VP, C.c_size_t, P(BatchPlan)],
    "asc_sparse_hip_batch_spmv": [_VP, P(BatchPlan)],
    "asc_sparse_hip_batch_graph_create": [_VP, P(BatchPlan), P(_VP)],
    "asc_sparse_hip_graph_launch": [_VP, _VP],
    "asc_sparse_hip_graph_destroy": [_VP],
    "asc_sparse_hip_batch_tune": [_VP, C.c_int64, P(BatchEntry), _VP, C.c_size_t, C.c_int32, P(BatchClasses), P(C.c_float)],
}

_lib = None


def load() -> C.CDLL:
    """Loads the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = Status
    lib.asc_sparse_hip_last_message.argtypes = []
    lib.asc_sparse_hip_last_message.restype = C.c_char_p
    lib.asc_sparse_hip_device_count.argtypes = []
    lib.asc_sparse_hip_device_count.restype = C.c_int32
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS) + ["asc_sparse_hip_last_message", "asc_sparse_hip_device_count"]


def check(st: Status) -> None:
    if st.code != OK:
        msg = load().asc_sparse_hip_last_message().decode()
        raise AscHipError(st.code, msg, st.native)


def device_count() -> int:
    return int(load().asc_sparse_hip_device_count())


class Context:
    """An asc_hip_context on one device (device=-1: validation only)."""

    def __init__(self, device: int = 0, stream: int | None = None,
                 determinism: int = DETERMINISTIC):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.asc_sparse_hip_context_create(device, stream, determinism, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            check(self.lib.asc_sparse_hip_context_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(self.lib.asc_sparse_hip_context_synchronize(self.handle))

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        check(self.lib.asc_sparse_hip_context_stream(self.handle, C.byref(s)))
        return s.value or 0

    # ---- memory --------------------------------------------------------------
    def alloc(self, nbytes: int, space: int = DEVICE) -> "Buffer":
        p = C.c_void_p()
        check(self.lib.asc_sparse_hip_allocate(self.handle, nbytes, space, C.byref(p)))
        return Buffer(self, p.value or 0, nbytes, space)

    def upload(self, arr: np.ndarray) -> "Buffer":
        arr = np.ascontiguousarray(arr)
        buf = self.alloc(max(arr.nbytes, 1))
        if arr.nbytes:
            check(self.lib.asc_sparse_hip_copy(self.handle, buf.ptr, arr.ctypes.data, arr.nbytes))
            self.synchronize()
        buf.dtype, buf.shape = arr.dtype, arr.shape
        return buf

    def download(self, buf: "Buffer", dtype, count: int, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(self.lib.asc_sparse_hip_copy(self.handle, out.ctypes.data,
                                               buf.ptr + offset_bytes, out.nbytes))
            self.synchronize()
        return out

    def copy(self, dst: int, src: int, nbytes: int):
        check(self.lib.asc_sparse_hip_copy(self.handle, dst, src, nbytes))

    def fill(self, buf: "Buffer", byte: int = 0, nbytes: int | None = None):
        check(self.lib.asc_sparse_hip_fill(self.handle, buf.ptr, byte,
                                           buf.nbytes if nbytes is None else nbytes))

    # ---- events --------------------------------------------------------------
    def record(self) -> "Event":
        e = C.c_void_p()
        check(self.lib.asc_sparse_hip_event_record(self.handle, C.byref(e)))
        return Event(self.lib, e)


class Event:
    def __init__(self, lib, handle):
        self.lib, self.handle = lib, handle

    def wait(self):
        check(self.lib.asc_sparse_hip_event_wait(self.handle))

    def elapsed_ms(self, later: "Event") -> float:
        v = C.c_float()
        check(self.lib.asc_sparse_hip_event_elapsed_ms(self.handle, later.handle, C.byref(v)))
        return float(v.value)

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.asc_sparse_hip_event_destroy(self.handle)
            self.handle = None


class Buffer:
    def __init__(self, ctx: Context, ptr: int, nbytes: int, space: int):
        self.ctx, self.ptr, self.nbytes, self.space = ctx, ptr, nbytes, space
        self.dtype, self.shape = None, None

    def free(self):
        if self.ptr and self.ctx.handle:
            check(self.ctx.lib.asc_sparse_hip_deallocate(self.ctx.handle, self.ptr, self.space))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def numpy(self, dtype=None, count=None) -> np.ndarray:
        dtype = dtype or self.dtype
        count = count if count is not None else self.nbytes // np.dtype(dtype).itemsize
        return self.ctx.download(self, dtype, count)


# ---- descriptor helpers ----------------------------------------------------------

def vector(buf_or_ptr, extent: int, kind: int, stride: int = 1, writable: bool = True,
           space: int = DEVICE) -> VectorDesc:
    ptr = buf_or_ptr.ptr if isinstance(buf_or_ptr, Buffer) else int(buf_or_ptr or 0)
    return VectorDesc(ptr, extent, stride, kind, space, 1 if writable else 0)


def dense(buf_or_ptr, rows: int, cols: int, ld: int, layout: int, kind: int,
          writable: bool = True, space: int = DEVICE) -> MatrixDesc:
    ptr = buf_or_ptr.ptr if isinstance(buf_or_ptr, Buffer) else int(buf_or_ptr or 0)
    return MatrixDesc(ptr, rows, cols, ld, layout, kind, space, 1 if writable else 0)


def host_sparse(fmt: int, rows: int, cols: int, offsets: np.ndarray, indices: np.ndarray | None,
                values: np.ndarray, kind: int) -> SparseDesc:
    """Descriptor over host numpy arrays (kept alive by the caller)."""
    nnz = int(values.shape[0])
    ib = indices.dtype.itemsize if indices is not None else 8
    return SparseDesc(offsets.ctypes.data, indices.ctypes.data if indices is not None else None,
                      values.ctypes.data, rows, cols, nnz, fmt, kind, ib, 8, HOST, 0)


@dataclass
class DeviceSparse:
    """Device CSR/CSC/COO arrays owned by Buffers plus their descriptor."""
    desc: SparseDesc
    buffers: list

    @property
    def rows(self):
        return self.desc.rows

    @property
    def cols(self):
        return self.desc.cols

    @property
    def nnz(self):
        return self.desc.nnz


def clone_compressed(ctx: Context, fmt: int, rows: int, cols: int, offsets: np.ndarray,
                     indices: np.ndarray, values: np.ndarray, kind: int,
                     index_bytes: int = 8) -> DeviceSparse:
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    indices = np.ascontiguousarray(indices)
    values = np.ascontiguousarray(values)
    src = host_sparse(fmt, rows, cols, offsets, indices, values, kind)
    outer = rows if fmt == CSR else cols
    nnz = values.shape[0]
    bo = ctx.alloc(8 * (outer + 1))
    bi = ctx.alloc(max(index_bytes * nnz, 1))
    bv = ctx.alloc(max(ELEM_BYTES[kind] * nnz, 1))
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_clone_compressed(ctx.handle, C.byref(src), index_bytes,
                                                  bo.ptr, bi.ptr, bv.ptr, C.byref(out)))
    return DeviceSparse(out, [bo, bi, bv])


def rebind_values(a: DeviceSparse, values) -> DeviceSparse:
    """The same trusted structure over other device values (RebindValues,
    compressed.h:472-516); `values` is a Buffer, kept alive by the result."""
    out = SparseDesc()
    ptr = values.ptr if hasattr(values, "ptr") else values
    check(load().asc_sparse_hip_rebind_values(C.byref(a.desc), ptr, C.byref(out)))
    r = DeviceSparse(out, [b for b in a.buffers] + ([values] if hasattr(values, "ptr") else []))
    if getattr(a, "analysis", None) is not None:
        r.analysis = a.analysis
    return r


def clone_coordinate(ctx: Context, rows: int, cols: int, coords: np.ndarray,
                     values: np.ndarray, kind: int) -> DeviceSparse:
    coords = np.ascontiguousarray(coords, dtype=np.int64).reshape(-1)
    values = np.ascontiguousarray(values)
    nnz = values.shape[0]
    src = SparseDesc(coords.ctypes.data, None, values.ctypes.data, rows, cols, nnz, COO, kind,
                     8, 8, HOST, 0)
    bc = ctx.alloc(max(16 * nnz, 1))
    bv = ctx.alloc(max(ELEM_BYTES[kind] * nnz, 1))
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_clone_coordinate(ctx.handle, C.byref(src), bc.ptr, bv.ptr,
                                                  C.byref(out)))
    return DeviceSparse(out, [bc, bv])


def narrow_indices(ctx: Context, a: DeviceSparse) -> DeviceSparse:
    b = ctx.alloc(max(4 * a.nnz, 1))
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_narrow_indices(ctx.handle, C.byref(a.desc), b.ptr, C.byref(out)))
    return DeviceSparse(out, a.buffers + [b])


def convert_values(ctx: Context, a: DeviceSparse, kind: int, scale: float = 1.0) -> DeviceSparse:
    b = ctx.alloc(max(ELEM_BYTES[kind] * a.nnz, 1))
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_convert_values_scaled(ctx.handle, C.byref(a.desc), kind, scale,
                                                       b.ptr, C.byref(out)))
    return DeviceSparse(out, a.buffers + [b])


def download_compressed(ctx: Context, a: DeviceSparse):
    d = a.desc
    outer = d.rows if d.format == CSR else d.cols
    off = ctx.download(_ptr_buf(ctx, d.outer_offsets), np.int64, outer + 1)
    idx = ctx.download(_ptr_buf(ctx, d.inner_indices), np.int32 if d.index_bytes == 4 else np.int64, d.nnz)
    val = ctx.download(_ptr_buf(ctx, d.values), NP_STORAGE[d.element_kind], d.nnz)
    return off, idx, val


class _PtrView:
    def __init__(self, ptr):
        self.ptr = ptr


def _ptr_buf(ctx, ptr):
    return _PtrView(ptr or 0)


def options(alg: int = ALG_AUTO, acc: int = ACC_PLAIN) -> SpmvOptions:
    return SpmvOptions(alg, acc)


class Workspace:
    """Grows-only device workspace reused across calls."""

    def __init__(self, ctx: Context):
        self.ctx, self.buf = ctx, None

    def get(self, nbytes: int):
        if nbytes == 0:
            return None, 0
        if self.buf is None or self.buf.nbytes < nbytes:
            if self.buf is not None:
                self.buf.free()
            self.buf = self.ctx.alloc(nbytes)
        return self.buf.ptr, self.buf.nbytes


def csr_spmv(ctx: Context, alpha: float, a: DeviceSparse, x: VectorDesc, beta: float,
             y: VectorDesc, alg: int = ALG_AUTO, acc: int = ACC_PLAIN,
             ws: Workspace | None = None):
    opt = options(alg, acc)
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_csr_spmv_workspace_size(ctx.handle, C.byref(a.desc), C.byref(x),
                                                         C.byref(y), C.byref(opt), C.byref(need)))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    check(ctx.lib.asc_sparse_hip_csr_spmv(ctx.handle, alpha, C.byref(a.desc), C.byref(x), beta,
                                          C.byref(y), p, n, C.byref(opt)))
    return ws


NOALLOC_SCRATCH_BYTES = 32 << 20


def csr_spmv_noalloc(ctx: Context, alpha: float, a: DeviceSparse, x: VectorDesc, beta: float,
                     y: VectorDesc, alg: int = ALG_AUTO, acc: int = ACC_PLAIN):
    """The allocation-free compat SpMV (the library's own scratch, no workspace)."""
    opt = options(alg, acc)
    check(ctx.lib.asc_sparse_hip_csr_spmv_noalloc(ctx.handle, alpha, C.byref(a.desc), C.byref(x),
                                                  beta, C.byref(y), C.byref(opt)))


def standard_spmv(ctx: Context, transpose: int, alpha: float, a: DeviceSparse, x: VectorDesc,
                  y: VectorDesc, alg: int = ALG_AUTO, ws: Workspace | None = None):
    opt = options(alg)
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_standard_spmv_workspace_size(
        ctx.handle, transpose, C.byref(a.desc), C.byref(x), C.byref(y), C.byref(opt),
        C.byref(need)))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    check(ctx.lib.asc_sparse_hip_standard_spmv(ctx.handle, transpose, Scalar(alpha, 0.0),
                                               C.byref(a.desc), C.byref(x), C.byref(y), p, n,
                                               C.byref(opt)))
    return ws


def standard_spmm(ctx: Context, transpose: int, alpha: float, a: DeviceSparse, b: MatrixDesc,
                  c: MatrixDesc, ws: Workspace | None = None, alg: int = ALG_AUTO):
    """C += alpha*op(A)*B; alg ALG_AUTO (plain form; the panel form when the
    context holds a stream-staged analysis of A listing only long rows),
    ALG_PANEL (plain form over the context's stream-staged column-panel
    analysis of A) or ALG_ORDERED (bit-exact)."""
    opt = options(alg)
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
        ctx.handle, transpose, C.byref(a.desc), C.byref(b), C.byref(c), C.byref(opt),
        C.byref(need)))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    check(ctx.lib.asc_sparse_hip_standard_spmm(ctx.handle, transpose, Scalar(alpha, 0.0),
                                               C.byref(a.desc), C.byref(b), C.byref(c), p, n,
                                               C.byref(opt)))
    return ws


def _output_buffers(ctx: Context, into: DeviceSparse | None, sizes: list) -> list:
    """Fresh output buffers of `sizes` bytes, or the (offsets, indices,
    values) buffers of an earlier result `into` when each is large enough."""
    if into is None:
        return [ctx.alloc(n) for n in sizes]
    bufs = into.buffers[:3]
    if len(bufs) != 3 or any(b.nbytes < n for b, n in zip(bufs, sizes)):
        raise ValueError("the output buffers to reuse are smaller than the conversion's output")
    return bufs


def convert_coordinate(ctx: Context, coo: DeviceSparse, out_format: int, index_bytes: int = 8,
                       ws: Workspace | None = None, into: DeviceSparse | None = None) -> DeviceSparse:
    """COO -> CSR/CSC into fresh device buffers, or into those of an earlier
    result `into` (no allocation: the C ABI's caller-provided outputs)."""
    d = coo.desc
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_convert_coordinate_workspace_size(
        ctx.handle, C.byref(d), out_format, index_bytes, C.byref(need)))
    outer = d.rows if out_format == CSR else d.cols
    bo, bi, bv = _output_buffers(ctx, into, [8 * (outer + 1), max(index_bytes * d.nnz, 1),
                                             max(ELEM_BYTES[d.element_kind] * d.nnz, 1)])
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_convert_coordinate(ctx.handle, C.byref(d), out_format,
                                                    index_bytes, bo.ptr, bi.ptr, bv.ptr,
                                                    C.byref(out), p, n))
    return DeviceSparse(out, [bo, bi, bv])


def convert_compressed(ctx: Context, a: DeviceSparse, ws: Workspace | None = None,
                       into: DeviceSparse | None = None) -> DeviceSparse:
    """CSR <-> CSC into fresh device buffers, or into those of an earlier
    result `into` (no allocation: the C ABI's caller-provided outputs)."""
    d = a.desc
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_convert_compressed_workspace_size(ctx.handle, C.byref(d),
                                                                   C.byref(need)))
    outer = d.cols if d.format == CSR else d.rows
    bo, bi, bv = _output_buffers(ctx, into, [8 * (outer + 1), max(d.index_bytes * d.nnz, 1),
                                             max(ELEM_BYTES[d.element_kind] * d.nnz, 1)])
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_convert_compressed(ctx.handle, C.byref(d), bo.ptr, bi.ptr,
                                                    bv.ptr, C.byref(out), p, n))
    return DeviceSparse(out, [bo, bi, bv])


def test_inject(what: int, countdown: int):
    check(load().asc_sparse_hip_test_inject(what, countdown))


# ---- batched SpMV -------------------------------------------------------------------

class Batch:
    """A batched-SpMV plan over device matrices and vectors (kept alive here)."""

    def __init__(self, ctx: Context, entries: list, keep: list | None = None,
                 lanes: list | None = None):
        self.ctx = ctx
        self.keep = keep or []
        self.n = len(entries)
        self.entries = (BatchEntry * max(self.n, 1))(*entries)
        need = C.c_size_t()
        check(ctx.lib.asc_sparse_hip_batch_plan_size(ctx.handle, self.n, self.entries, C.byref(need)))
        self.mem = ctx.alloc(max(need.value, 256))
        self.plan = BatchPlan()
        self.graph = None
        self.build(lanes)

    def build(self, lanes=None):
        cls = None
        if lanes is not None:
            cls = BatchClasses((C.c_uint8 * 8)(*lanes))
        check(self.ctx.lib.asc_sparse_hip_batch_plan_build(
            self.ctx.handle, self.n, self.entries, C.byref(cls) if cls else None, self.mem.ptr,
            self.mem.nbytes, C.byref(self.plan)))
        self.drop_graph()

    def tune(self, reps: int = 3):
        out = BatchClasses()
        ms = (C.c_float * 64)()
        check(self.ctx.lib.asc_sparse_hip_batch_tune(self.ctx.handle, self.n, self.entries, self.mem.ptr,
                                                     self.mem.nbytes, reps, C.byref(out), ms))
        lanes = list(out.lanes)
        self.build(lanes)
        return lanes, [list(ms[c * 8:(c + 1) * 8]) for c in range(8)]

    def run(self):
        check(self.ctx.lib.asc_sparse_hip_batch_spmv(self.ctx.handle, C.byref(self.plan)))

    def capture(self):
        self.drop_graph()
        g = C.c_void_p()
        check(self.ctx.lib.asc_sparse_hip_batch_graph_create(self.ctx.handle, C.byref(self.plan), C.byref(g)))
        self.graph = g

    def replay(self):
        check(self.ctx.lib.asc_sparse_hip_graph_launch(self.ctx.handle, self.graph))

    def drop_graph(self):
        if self.graph:
            check(self.ctx.lib.asc_sparse_hip_graph_destroy(self.graph))
            self.graph = None

    def __del__(self):
        try:
            self.drop_graph()
        except Exception:
            pass


def upload_packed(ctx: Context, packed, x: np.ndarray, y: np.ndarray, kind: int,
                  alpha: float = 1.0, beta: float = 0.0):
    """Uploads a packed batch (see gen.pack) and returns (entries, buffers).

    `x` and `y` are the concatenated per-matrix vectors (square matrices)."""
    dt = np.float64 if kind == F64 else np.float32
    bo = ctx.upload(packed.offsets)
    bi = ctx.upload(packed.indices)
    bv = ctx.upload(packed.values.astype(dt))
    bx = ctx.upload(x.astype(dt))
    by = ctx.upload(y.astype(dt))
    ib = packed.indices.dtype.itemsize
    es = ELEM_BYTES[kind]
    entries = []
    for i in range(packed.rows.shape[0]):
        r, n = int(packed.rows[i]), int(packed.nnz[i])
        a = SparseDesc(bo.ptr + 8 * int(packed.off_start[i]), bi.ptr + ib * int(packed.nz_start[i]),
                       bv.ptr + es * int(packed.nz_start[i]), r, r, n, CSR, kind, ib, 8, DEVICE, 1)
        vs = int(packed.vec_start[i])
        entries.append(BatchEntry(a, VectorDesc(bx.ptr + es * vs, r, 1, kind, DEVICE, 0),
                                  VectorDesc(by.ptr + es * vs, r, 1, kind, DEVICE, 1), alpha, beta))
    return entries, [bo, bi, bv, bx, by]


# ---- variant selection: autotuner, plan cache, features, MLP --------------------------

class Selector(C.Structure):
    _fields_ = [("device_model", C.c_void_p), ("outputs", C.c_int32), ("margin", C.c_float),
                ("clip", C.c_float), ("reserved", C.c_int32)]


class Selection(C.Structure):
    _fields_ = [("algorithm", C.c_uint8), ("staging", C.c_uint8), ("direct_algorithm", C.c_uint8),
                ("fell_back", C.c_uint8), ("reserved", C.c_uint8 * 4)]


SELECT_AMORTIZE = 100  # SpMVs over which a plan's analysis time is charged
SELECT_DIRECT_OUTPUTS, SELECT_PLAN_OUTPUTS = 8, 11


_SIGS2 = {
    "asc_sparse_hip_plan_set": [_VP, P(SparseDesc), C.c_uint8],
    "asc_sparse_hip_plan_get": [_VP, P(SparseDesc), P(C.c_uint8)],
    "asc_sparse_hip_plan_clear": [_VP],
    "asc_sparse_hip_csr_spmv_autotune": [_VP, P(SparseDesc), P(VectorDesc), P(VectorDesc), _VP, C.c_size_t,
                                         C.c_int32, P(C.c_float), P(C.c_uint8)],
    "asc_sparse_hip_csr_features_workspace_size": [_VP, P(C.c_size_t)],
    "asc_sparse_hip_csr_features": [_VP, P(SparseDesc), _VP, _VP, C.c_size_t],
    "asc_sparse_hip_mlp_sizes": [C.c_int32, C.c_int64, P(C.c_size_t), P(C.c_size_t)],
    "asc_sparse_hip_mlp_forward": [_VP, C.c_int32, _VP, C.c_int64, _VP, _VP, _VP, C.c_size_t],
    "asc_sparse_hip_mlp_gradient": [_VP, C.c_int32, _VP, C.c_int64, _VP, _VP, _VP, _VP, _VP, C.c_size_t],
    "asc_sparse_hip_mlp_adam": [_VP, _VP, _VP, _VP, _VP, C.c_int64, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_int64],
    "asc_sparse_hip_normalize_features": [_VP, _VP, _VP, _VP, C.c_int64, _VP],
    "asc_sparse_hip_select_workspace_size": [_VP, P(Selector), P(C.c_size_t)],
    "asc_sparse_hip_select": [_VP, P(SparseDesc), P(Selector), _VP, C.c_size_t, P(Selection), P(C.c_float)],
}
_SIGS.update(_SIGS2)
if _lib is not None:  # signatures added after an early load()
    for _n, _a in _SIGS2.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


def plan_set(ctx: Context, a: DeviceSparse, alg: int):
    check(ctx.lib.asc_sparse_hip_plan_set(ctx.handle, C.byref(a.desc), alg))


def plan_get(ctx: Context, a: DeviceSparse) -> int:
    v = C.c_uint8()
    check(ctx.lib.asc_sparse_hip_plan_get(ctx.handle, C.byref(a.desc), C.byref(v)))
    return v.value


def plan_clear(ctx: Context):
    check(ctx.lib.asc_sparse_hip_plan_clear(ctx.handle))


def autotune(ctx: Context, a: DeviceSparse, x: VectorDesc, y: VectorDesc, reps: int = 3,
             ws: Workspace | None = None):
    """Times every tunable variant; returns (best algorithm, {alg: ms})."""
    opt = options(ALG_MERGE)
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_csr_spmv_workspace_size(ctx.handle, C.byref(a.desc), C.byref(x),
                                                         C.byref(y), C.byref(opt), C.byref(need)))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    times = (C.c_float * 10)(*([-1.0] * 10))
    best = C.c_uint8()
    check(ctx.lib.asc_sparse_hip_csr_spmv_autotune(ctx.handle, C.byref(a.desc), C.byref(x), C.byref(y),
                                                   p, n, reps, times, C.byref(best)))
    return best.value, {alg: float(times[alg]) for alg in TUNABLE_ALGS}


def features(ctx: Context, a: DeviceSparse, ws: Workspace | None = None) -> np.ndarray:
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_csr_features_workspace_size(ctx.handle, C.byref(need)))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    out = ctx.alloc(64)
    check(ctx.lib.asc_sparse_hip_csr_features(ctx.handle, C.byref(a.desc), out.ptr, p, n))
    return ctx.download(out, np.float32, 16)


def mlp_sizes(k: int, batch: int):
    pf, sb = C.c_size_t(), C.c_size_t()
    check(load().asc_sparse_hip_mlp_sizes(k, batch, C.byref(pf), C.byref(sb)))
    return pf.value, sb.value

_SIGS3 = {
    "asc_sparse_hip_convert_to_coordinate_workspace_size": [_VP, P(SparseDesc), P(C.c_size_t)],
    "asc_sparse_hip_convert_to_coordinate": [_VP, P(SparseDesc), _VP, _VP, P(SparseDesc), _VP, C.c_size_t],
}
_SIGS.update(_SIGS3)
if _lib is not None:
    for _n, _a in _SIGS3.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


def convert_to_coordinate(ctx: Context, a: DeviceSparse, ws: Workspace | None = None) -> DeviceSparse:
    d = a.desc
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_convert_to_coordinate_workspace_size(ctx.handle, C.byref(d), C.byref(need)))
    bc = ctx.alloc(max(16 * d.nnz, 1))
    bv = ctx.alloc(max(ELEM_BYTES[d.element_kind] * d.nnz, 1))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_convert_to_coordinate(ctx.handle, C.byref(d), bc.ptr, bv.ptr,
                                                       C.byref(out), p, n))
    return DeviceSparse(out, [bc, bv])


STAGING_AUTO, STAGING_L2, STAGING_LDS, STAGING_STREAM = 0, 1, 2, 3


class PanelOptions(C.Structure):
    _fields_ = [("panel_columns", C.c_int64), ("long_row_min", C.c_int64),
                ("short_algorithm", C.c_uint8), ("long_algorithm", C.c_uint8),
                ("staging", C.c_uint8), ("reserved", C.c_uint8 * 5)]


_SIGS4 = {
    "asc_sparse_hip_csr_spmv_analysis_size": [_VP, P(SparseDesc), P(PanelOptions), P(C.c_size_t)],
    "asc_sparse_hip_csr_spmv_analyze": [_VP, P(SparseDesc), P(PanelOptions), _VP, C.c_size_t],
    "asc_sparse_hip_csr_spmv_refresh_values": [_VP, P(SparseDesc)],
}
_SIGS.update(_SIGS4)
if _lib is not None:
    for _n, _a in _SIGS4.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


def panel_options(panel_columns: int = 0, long_row_min: int = 0, short_alg: int = ALG_AUTO,
                  long_alg: int = ALG_AUTO, staging: int = STAGING_AUTO) -> PanelOptions:
    return PanelOptions(panel_columns, long_row_min, short_alg, long_alg, staging)


def csr_spmv_analysis_size(ctx: Context, a: DeviceSparse, opts: PanelOptions | None = None) -> int:
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_csr_spmv_analysis_size(ctx.handle, C.byref(a.desc),
                                                        C.byref(opts) if opts else None, C.byref(need)))
    return need.value


def csr_spmv_analyze(ctx: Context, a: DeviceSparse, opts: PanelOptions | None = None) -> Buffer:
    """Column-panel analysis: registers ALG_PANEL for `a` in the context's plan
    cache.  The analysis buffer is kept alive with the matrix (`a.analysis`),
    since the plan reads it on every ALG_PANEL / ALG_AUTO call."""
    need = csr_spmv_analysis_size(ctx, a, opts)
    buffer = ctx.alloc(need)
    check(ctx.lib.asc_sparse_hip_csr_spmv_analyze(ctx.handle, C.byref(a.desc),
                                                  C.byref(opts) if opts else None, buffer.ptr,
                                                  buffer.nbytes))
    previous = getattr(a, "analysis", None)
    a.analysis = buffer
    if previous is not None:
        ctx.synchronize()  # queued work may still read the replaced plan
        previous.free()
    return buffer


def csr_spmv_refresh_values(ctx: Context, a: DeviceSparse):
    """Copies a's current values into its stream-staged panel analysis (the
    values-only refresh after an in-place change or a rebind)."""
    check(ctx.lib.asc_sparse_hip_csr_spmv_refresh_values(ctx.handle, C.byref(a.desc)))


DUPLICATE_REJECT, DUPLICATE_SUM = 0, 1
ZERO_KEEP, ZERO_DROP = 0, 1

_SIGS5 = {
    "asc_sparse_hip_coo_finalize_workspace_size": [_VP, P(SparseDesc), P(C.c_size_t)],
    "asc_sparse_hip_coo_finalize": [_VP, P(SparseDesc), C.c_uint8, C.c_uint8, _VP, _VP, P(SparseDesc),
                                    _VP, C.c_size_t],
}
_SIGS.update(_SIGS5)
if _lib is not None:
    for _n, _a in _SIGS5.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


def triplets(ctx: Context, rows: int, cols: int, coords: np.ndarray, values: np.ndarray,
             kind: int) -> DeviceSparse:
    """Raw (row, col, value) triplets in insertion order, uploaded as an
    untrusted device COO descriptor for coo_finalize."""
    coords = np.ascontiguousarray(coords, dtype=np.int64).reshape(-1)
    vals = np.ascontiguousarray(values, dtype=NP_STORAGE[kind])
    nnz = vals.shape[0]
    bc = ctx.upload(coords) if nnz else ctx.alloc(16)
    bv = ctx.upload(vals) if nnz else ctx.alloc(8)
    d = SparseDesc(bc.ptr, None, bv.ptr, rows, cols, nnz, COO, kind, 8, 8, DEVICE, 0)
    return DeviceSparse(d, [bc, bv])


def coo_finalize(ctx: Context, t: DeviceSparse, duplicate_policy: int, zero_policy: int,
                 ws: Workspace | None = None) -> DeviceSparse:
    """Device CoordinateBuilder::Finalize: canonical COO from raw triplets."""
    d = t.desc
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_coo_finalize_workspace_size(ctx.handle, C.byref(d), C.byref(need)))
    bc = ctx.alloc(max(16 * d.nnz, 16))
    bv = ctx.alloc(max(ELEM_BYTES[d.element_kind] * d.nnz, 8))
    ws = ws or Workspace(ctx)
    p, n = ws.get(need.value)
    out = SparseDesc()
    check(ctx.lib.asc_sparse_hip_coo_finalize(ctx.handle, C.byref(d), duplicate_policy, zero_policy,
                                              bc.ptr, bv.ptr, C.byref(out), p, n))
    return DeviceSparse(out, [bc, bv])


LEVEL1_DOT, LEVEL1_AXPY, LEVEL1_GATHER, LEVEL1_GATHER_ZERO, LEVEL1_SCATTER = 0, 1, 2, 3, 4
UNCONJUGATED, CONJUGATED = 0, 1


class IndexedDesc(C.Structure):
    _fields_ = [("indices", C.c_void_p), ("values", C.c_void_p), ("nnz", C.c_int64),
                ("dense_extent", C.c_int64), ("element_kind", C.c_uint8), ("memory_space", C.c_uint8),
                ("writable", C.c_uint8), ("trusted", C.c_uint8), ("reserved", C.c_uint8 * 4)]


_SIGS7 = {
    "asc_sparse_hip_clone_indexed": [_VP, P(IndexedDesc), _VP, _VP, P(IndexedDesc)],
    "asc_sparse_hip_standard_level1": [_VP, C.c_uint8, C.c_uint8, Scalar, P(IndexedDesc),
                                       P(VectorDesc), P(VectorDesc), P(SpmvOptions)],
}
_SIGS.update(_SIGS7)
if _lib is not None:
    for _n, _a in _SIGS7.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


class DeviceIndexed:
    """A trusted device indexed vector and the buffers that own it."""

    def __init__(self, desc: IndexedDesc, buffers: list):
        self.desc, self.buffers = desc, buffers


def clone_indexed(ctx: Context, indices: np.ndarray, values: np.ndarray, dense_extent: int,
                  kind: int) -> DeviceIndexed:
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    vals = np.ascontiguousarray(values, dtype=NP_STORAGE[kind])
    src = IndexedDesc(idx.ctypes.data, vals.ctypes.data, idx.shape[0], dense_extent, kind, HOST, 0, 0)
    bi = ctx.alloc(max(8 * idx.shape[0], 8))
    bv = ctx.alloc(max(vals.nbytes, 8))
    out = IndexedDesc()
    check(ctx.lib.asc_sparse_hip_clone_indexed(ctx.handle, C.byref(src), bi.ptr, bv.ptr, C.byref(out)))
    return DeviceIndexed(out, [bi, bv])


def level1(ctx: Context, op: int, sparse: DeviceIndexed, dense: VectorDesc, alpha: float = 1.0,
           result: VectorDesc | None = None, alg: int = ALG_AUTO, conjugation: int = UNCONJUGATED):
    opt = options(alg)
    check(ctx.lib.asc_sparse_hip_standard_level1(ctx.handle, op, conjugation, Scalar(alpha, 0.0),
                                                 C.byref(sparse.desc), C.byref(dense),
                                                 C.byref(result) if result is not None else None,
                                                 C.byref(opt)))


UPPER, LOWER = 0, 1
NON_UNIT, UNIT = 0, 1


class TriangularInfo(C.Structure):
    _fields_ = [("rows", C.c_int64), ("nnz", C.c_int64), ("levels", C.c_int64),
                ("max_level_rows", C.c_int64), ("offsets", C.c_void_p), ("indices", C.c_void_p),
                ("values", C.c_void_p), ("transpose", C.c_uint8), ("triangle", C.c_uint8),
                ("diagonal", C.c_uint8), ("format", C.c_uint8), ("reserved", C.c_uint8 * 4)]


_SIGS8 = {
    "asc_sparse_hip_triangular_analysis_size": [_VP, C.c_uint8, P(SparseDesc), P(C.c_size_t)],
    "asc_sparse_hip_triangular_analyze": [_VP, C.c_uint8, C.c_uint8, C.c_uint8, P(SparseDesc), _VP,
                                          C.c_size_t, P(TriangularInfo)],
    "asc_sparse_hip_triangular_solve": [_VP, C.c_uint8, Scalar, P(SparseDesc), P(MatrixDesc), _VP,
                                        C.c_size_t, P(TriangularInfo)],
}
_SIGS.update(_SIGS8)
if _lib is not None:
    for _n, _a in _SIGS8.items():
        getattr(_lib, _n).argtypes = _a
        getattr(_lib, _n).restype = Status


class TriangularPlan:
    """The analysis buffer and its info for one matrix and operation."""

    def __init__(self, buffer: Buffer, info: TriangularInfo):
        self.buffer, self.info = buffer, info


def triangular_analyze(ctx: Context, transpose: int, triangle: int, diagonal: int,
                       a: DeviceSparse) -> TriangularPlan:
    need = C.c_size_t()
    check(ctx.lib.asc_sparse_hip_triangular_analysis_size(ctx.handle, transpose, C.byref(a.desc),
                                                          C.byref(need)))
    buf = ctx.alloc(need.value)
    info = TriangularInfo()
    check(ctx.lib.asc_sparse_hip_triangular_analyze(ctx.handle, transpose, triangle, diagonal,
                                                    C.byref(a.desc), buf.ptr, buf.nbytes,
                                                    C.byref(info)))
    return TriangularPlan(buf, info)


def triangular_solve(ctx: Context, transpose: int, alpha: float, a: DeviceSparse, rhs: MatrixDesc,
                     plan: TriangularPlan):
    check(ctx.lib.asc_sparse_hip_triangular_solve(ctx.handle, transpose, Scalar(alpha, 0.0),
                                                  C.byref(a.desc), C.byref(rhs), plan.buffer.ptr,
                                                  plan.buffer.nbytes, C.byref(plan.info)))
