"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

Only tests, the smoke check and the benchmark's CPU-baseline leg load the
oracle, and only as a checker.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_driver.so")

_lib = None
I64, F64P, F32P, VP = C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            raise ImportError("oracle/liboracle.so is missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE)
        L.oracle_validate_compressed.argtypes = [I64, I64, VP, VP, I64]
        L.oracle_compat_spmv_f64.argtypes = [I64, VP, VP, VP, C.c_double, VP, I64, C.c_double, VP, I64]
        L.oracle_compat_spmv_f32.argtypes = [I64, VP, VP, VP, C.c_float, VP, I64, C.c_float, VP, I64]
        L.oracle_std_spmv_f64.argtypes = [C.c_int, C.c_int, I64, I64, VP, VP, VP, C.c_double, VP, I64, VP, I64]
        L.oracle_std_spmv_f32.argtypes = [C.c_int, C.c_int, I64, I64, VP, VP, VP, C.c_float, VP, I64, VP, I64]
        L.oracle_std_spmm_f64.argtypes = [C.c_int, C.c_int, I64, I64, VP, VP, VP, C.c_double, VP, C.c_int, I64, I64, VP, C.c_int, I64]
        L.oracle_std_spmm_f32.argtypes = [C.c_int, C.c_int, I64, I64, VP, VP, VP, C.c_float, VP, C.c_int, I64, I64, VP, C.c_int, I64]
        L.oracle_abs_spmv_f64.argtypes = [I64, VP, VP, VP, VP, VP]
        L.oracle_abs_spmv_f64.restype = None
        L.oracle_abs_spmv_f64_i32.argtypes = [I64, VP, VP, VP, VP, VP]
        L.oracle_abs_spmv_f64_i32.restype = None
        L.oracle_compat_spmv_f64_i32.argtypes = L.oracle_compat_spmv_f64.argtypes
        L.oracle_compat_spmv_f32_i32.argtypes = L.oracle_compat_spmv_f32.argtypes
        L.oracle_std_spmm_f64_i32.argtypes = L.oracle_std_spmm_f64.argtypes
        L.oracle_coo_to_compressed_faithful.argtypes = [C.c_int, I64, I64, I64, VP, VP, C.c_int, VP, VP, VP]
        L.oracle_coo_to_compressed_fast.argtypes = [C.c_int, I64, I64, I64, VP, VP, C.c_int, VP, VP, VP]
        L.oracle_transpose_faithful.argtypes = [C.c_int, I64, I64, VP, VP, VP, C.c_int, VP, VP, VP]
        L.oracle_transpose_fast.argtypes = [C.c_int, I64, I64, VP, VP, VP, C.c_int, VP, VP, VP]
        L.oracle_compressed_to_coo.argtypes = [C.c_int, I64, VP, VP, VP, C.c_int, VP, VP]
        L.oracle_coo_finalize_f64.argtypes = [I64, VP, VP, C.c_int, C.c_int, VP, VP, VP]
        L.oracle_coo_finalize_f32.argtypes = [I64, VP, VP, C.c_int, C.c_int, VP, VP, VP]
        for sfx, ct in (("f64", C.c_double), ("f32", C.c_float)):
            getattr(L, f"oracle_sparse_dot_{sfx}").argtypes = [I64, VP, VP, VP, I64]
            getattr(L, f"oracle_sparse_dot_{sfx}").restype = ct
            getattr(L, f"oracle_sparse_axpy_{sfx}").argtypes = [I64, ct, VP, VP, VP, I64]
            for op in ("gather", "gather_zero", "scatter"):
                getattr(L, f"oracle_sparse_{op}_{sfx}").argtypes = [I64, VP, VP, VP, I64]
            getattr(L, f"oracle_trsv_{sfx}").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, I64, VP,
                                                         VP, VP, ct, VP, I64, I64, I64]
        _lib = L
    return _lib


def _p(a):
    if a is not None and a.size and not a.flags.c_contiguous:
        raise ValueError("oracle arrays must be contiguous (pass strides as incx/incy)")
    return a.ctypes.data if a is not None and a.size else None


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _idx(indices):
    """int32 indices pass as they are (the *_i32 entry points), any other
    integer type as int64."""
    indices = np.asarray(indices)
    if indices.dtype == np.int32 and indices.flags.c_contiguous:
        return indices, True
    return i64(indices), False


def compat_spmv(offsets, indices, values, alpha, x, beta, y, incx=1, incy=1):
    """y (modified in place) = alpha*A*x (+ beta*y when beta != 0).  Row r
    reads indices[offsets[r]:offsets[r+1]], so `offsets` may be any row
    range's slice of the full offsets over the full indices and values."""
    offsets = i64(offsets)
    indices, narrow = _idx(indices)
    rows = offsets.shape[0] - 1
    sfx = ("f64" if values.dtype == np.float64 else "f32") + ("_i32" if narrow else "")
    getattr(lib(), "oracle_compat_spmv_" + sfx)(rows, _p(offsets), _p(indices), _p(values), alpha,
                                                _p(x), incx, beta, _p(y), incy)
    return y


def std_spmv(fmt, op, rows, cols, offsets, indices, values, alpha, x, y, incx=1, incy=1):
    offsets, indices = i64(offsets), i64(indices)
    f = lib().oracle_std_spmv_f64 if values.dtype == np.float64 else lib().oracle_std_spmv_f32
    f(fmt, op, rows, cols, _p(offsets), _p(indices), _p(values), alpha, _p(x), incx, _p(y), incy)
    return y


def std_spmm(fmt, op, rows, cols, offsets, indices, values, alpha, b, b_layout, ldb, k, c, c_layout, ldc):
    offsets, indices = i64(offsets), np.asarray(indices)
    if indices.dtype == np.int32 and values.dtype == np.float64:
        indices, _ = _idx(indices)
        f = lib().oracle_std_spmm_f64_i32
    else:
        indices = i64(indices)
        f = lib().oracle_std_spmm_f64 if values.dtype == np.float64 else lib().oracle_std_spmm_f32
    f(fmt, op, rows, cols, _p(offsets), _p(indices), _p(values), alpha, _p(b), b_layout, ldb, k,
      _p(c), c_layout, ldc)
    return c


def abs_spmv(offsets, indices, values, x, out=None):
    """(|A||x|)_i per row (FP64 values and x); `offsets` may be a row
    range's slice, as for compat_spmv."""
    offsets = i64(offsets)
    indices, narrow = _idx(indices)
    if out is None:
        out = np.empty(offsets.shape[0] - 1)
    f = lib().oracle_abs_spmv_f64_i32 if narrow else lib().oracle_abs_spmv_f64
    f(out.shape[0], _p(offsets), _p(indices), _p(values), _p(x), _p(out))
    return out


def row_chunks(offsets, parts):
    """Row ranges [r0, r1) of about equal entry counts (for running the
    serial oracle over independent rows on several host threads)."""
    rows = offsets.shape[0] - 1
    nnz = int(offsets[-1])
    cuts = np.searchsorted(offsets, np.linspace(0, nnz, parts + 1)[1:-1]).tolist()
    bounds = sorted(set([0] + [min(max(c, 0), rows) for c in cuts] + [rows]))
    return list(zip(bounds[:-1], bounds[1:]))


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def parallel_rows(offsets, fn, parts=None):
    """Calls fn(r0, r1) for row chunks on a thread pool (the oracle's C
    loops release the GIL), and re-raises the first failure."""
    from concurrent.futures import ThreadPoolExecutor
    n = parts or 4 * threads()
    with ThreadPoolExecutor(threads()) as ex:
        for f in [ex.submit(fn, r0, r1) for r0, r1 in row_chunks(offsets, n)]:
            f.result()


def compat_spmv_rows(offsets, indices, values, alpha, x, beta, y):
    """compat_spmv over every row, split over host threads (unit strides)."""
    def part(r0, r1):
        compat_spmv(offsets[r0:r1 + 1], indices, values, alpha, x, beta, y[r0:r1])
    parallel_rows(offsets, part)
    return y


def abs_spmv_rows(offsets, indices, values, x):
    """abs_spmv over every row, split over host threads."""
    out = np.empty(offsets.shape[0] - 1)

    def part(r0, r1):
        abs_spmv(offsets[r0:r1 + 1], indices, values, x, out[r0:r1])
    parallel_rows(offsets, part)
    return out


def coo_to_compressed(fmt, rows, cols, coords, values, faithful=False):
    coords = i64(coords).reshape(-1)
    nnz = values.shape[0]
    outer = rows if fmt == 0 else cols
    off = np.empty(outer + 1, dtype=np.int64)
    idx = np.empty(nnz, dtype=np.int64)
    val = np.empty_like(values)
    f = lib().oracle_coo_to_compressed_faithful if faithful else lib().oracle_coo_to_compressed_fast
    f(fmt, rows, cols, nnz, _p(coords), _p(values), values.itemsize, _p(off), _p(idx), _p(val))
    return off, idx, val


def transpose(src_fmt, rows, cols, offsets, indices, values, faithful=False):
    offsets, indices = i64(offsets), i64(indices)
    out_outer = cols if src_fmt == 0 else rows
    nnz = values.shape[0]
    off = np.empty(out_outer + 1, dtype=np.int64)
    idx = np.empty(nnz, dtype=np.int64)
    val = np.empty_like(values)
    f = lib().oracle_transpose_faithful if faithful else lib().oracle_transpose_fast
    f(src_fmt, rows, cols, _p(offsets), _p(indices), _p(values), values.itemsize, _p(off), _p(idx), _p(val))
    return off, idx, val


def compressed_to_coo(src_fmt, outer, offsets, indices, values):
    offsets, indices = i64(offsets), i64(indices)
    nnz = values.shape[0]
    coords = np.empty(2 * nnz, dtype=np.int64)
    val = np.empty_like(values)
    lib().oracle_compressed_to_coo(src_fmt, outer, _p(offsets), _p(indices), _p(values), values.itemsize,
                                   _p(coords), _p(val))
    return coords, val


def finalize(coords, values, duplicate_policy, zero_policy):
    """FP64 or FP32 by the dtype of `values` (anything else is taken as FP64)."""
    coords = i64(coords).reshape(-1)
    dt = np.float32 if np.asarray(values).dtype == np.float32 else np.float64
    values = np.ascontiguousarray(values, dtype=dt)
    nnz = values.shape[0]
    oc = np.empty(max(2 * nnz, 2), dtype=np.int64)
    ov = np.empty(max(nnz, 1), dtype=dt)
    n = C.c_int64()
    fn = lib().oracle_coo_finalize_f32 if dt == np.float32 else lib().oracle_coo_finalize_f64
    rc = fn(nnz, _p(coords), _p(values), duplicate_policy, zero_policy, _p(oc), _p(ov), C.byref(n))
    return rc, oc[: 2 * n.value], ov[: n.value]


def validate(outer, inner, offsets, indices, nnz):
    offsets, indices = i64(offsets), i64(indices)
    return lib().oracle_validate_compressed(outer, inner, _p(offsets), _p(indices), nnz)


# ---- the reference compiled from its own sources (oracle/_ref) ---------------------

_ref = None


def ref_available():
    return os.path.exists(REF)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF)
        R.ref_compat_spmv_f64.argtypes = [I64, I64, I64, VP, VP, VP, C.c_double, VP, I64, C.c_double, VP, I64]
        R.ref_compat_spmv_f32.argtypes = [I64, I64, I64, VP, VP, VP, C.c_float, VP, I64, C.c_float, VP, I64]
        R.ref_std_spmv_f64.argtypes = [C.c_int, C.c_int, I64, I64, I64, VP, VP, VP, C.c_double, VP, I64, VP, I64]
        R.ref_std_spmm_f64.argtypes = [C.c_int, C.c_int, I64, I64, I64, VP, VP, VP, C.c_double, VP, C.c_int, I64, I64, VP, C.c_int, I64]
        R.ref_coo_to_compressed_f64.argtypes = [C.c_int, I64, I64, I64, VP, VP, VP, VP, VP]
        R.ref_transpose_f64.argtypes = [C.c_int, I64, I64, VP, VP, VP, VP, VP, VP]
        R.ref_finalize_f64.argtypes = [I64, I64, I64, VP, VP, C.c_int, C.c_int, VP, VP, VP]
        R.ref_finalize_f32.argtypes = [I64, I64, I64, VP, VP, C.c_int, C.c_int, VP, VP, VP]
        R.ref_level1_f64.argtypes = [C.c_int, I64, I64, VP, VP, VP, I64, C.c_double, VP]
        R.ref_level1_f32.argtypes = [C.c_int, I64, I64, VP, VP, VP, I64, C.c_float, VP]
        R.ref_trsv_f64.argtypes = [C.c_int] * 4 + [I64, I64, VP, VP, VP, C.c_double, VP, I64, C.c_int]
        R.ref_trsv_f32.argtypes = [C.c_int] * 4 + [I64, I64, VP, VP, VP, C.c_float, VP, I64, C.c_int]
        _ref = R
    return _ref


def ref_compat_spmv(offsets, indices, values, alpha, x, beta, y, incx=1, incy=1, cols=None):
    offsets, indices = i64(offsets), i64(indices)
    rows = offsets.shape[0] - 1
    cols = (x.shape[0] + abs(incx) - 1) // abs(incx) if cols is None else cols
    f = ref().ref_compat_spmv_f64 if values.dtype == np.float64 else ref().ref_compat_spmv_f32
    rc = f(rows, cols, values.shape[0], _p(offsets), _p(indices), _p(values), alpha, _p(x), incx,
           beta, _p(y), incy)
    if rc != 0:
        raise RuntimeError(f"reference Spmv failed with ErrorCode {rc}")
    return y


def ref_std_spmv(fmt, op, rows, cols, offsets, indices, values, alpha, x, y):
    offsets, indices = i64(offsets), i64(indices)
    rc = ref().ref_std_spmv_f64(fmt, op, rows, cols, values.shape[0], _p(offsets), _p(indices),
                                _p(values), alpha, _p(x), 1, _p(y), 1)
    if rc != 0:
        raise RuntimeError(f"reference Spmv failed with ErrorCode {rc}")
    return y


def ref_std_spmm(fmt, op, rows, cols, offsets, indices, values, alpha, b, b_layout, ldb, k, c, c_layout, ldc):
    offsets, indices = i64(offsets), i64(indices)
    rc = ref().ref_std_spmm_f64(fmt, op, rows, cols, values.shape[0], _p(offsets), _p(indices),
                                _p(values), alpha, _p(b), b_layout, ldb, k, _p(c), c_layout, ldc)
    if rc != 0:
        raise RuntimeError(f"reference Spmm failed with ErrorCode {rc}")
    return c


def ref_coo_to_compressed(fmt, rows, cols, coords, values):
    coords = i64(coords).reshape(-1)
    nnz = values.shape[0]
    outer = rows if fmt == 0 else cols
    off = np.empty(outer + 1, np.int64)
    idx = np.empty(max(nnz, 1), np.int64)
    val = np.empty(max(nnz, 1))
    rc = ref().ref_coo_to_compressed_f64(fmt, rows, cols, nnz, _p(coords), _p(values), _p(off),
                                         _p(idx), _p(val))
    if rc != 0:
        raise RuntimeError(f"reference conversion failed with ErrorCode {rc}")
    return off, idx[:nnz], val[:nnz]


def ref_transpose(src_fmt, rows, cols, offsets, indices, values):
    offsets, indices = i64(offsets), i64(indices)
    nnz = values.shape[0]
    outer = cols if src_fmt == 0 else rows
    off = np.empty(outer + 1, np.int64)
    idx = np.empty(max(nnz, 1), np.int64)
    val = np.empty(max(nnz, 1))
    rc = ref().ref_transpose_f64(src_fmt, rows, cols, _p(offsets), _p(indices), _p(values),
                                 _p(off), _p(idx), _p(val))
    if rc != 0:
        raise RuntimeError(f"reference conversion failed with ErrorCode {rc}")
    return off, idx[:nnz], val[:nnz]


def ref_finalize(rows, cols, coords, values, duplicate_policy, zero_policy):
    coords = i64(coords).reshape(-1)
    dt = np.float32 if np.asarray(values).dtype == np.float32 else np.float64
    values = np.ascontiguousarray(values, dtype=dt)
    nnz = values.shape[0]
    oc = np.empty(max(2 * nnz, 2), np.int64)
    ov = np.empty(max(nnz, 1), dt)
    n = C.c_int64()
    fn = ref().ref_finalize_f32 if dt == np.float32 else ref().ref_finalize_f64
    rc = fn(rows, cols, nnz, _p(coords), _p(values), duplicate_policy, zero_policy, _p(oc),
            _p(ov), C.byref(n))
    return rc, oc[: 2 * n.value], ov[: n.value]


# ---- sparse Level-1 -------------------------------------------------------------
LEVEL1_OPS = ("dot", "axpy", "gather", "gather_zero", "scatter")


def level1(op, idx, v, x, inc=1, alpha=1.0):
    """Oracle Level-1 op on copies; returns (dot or None, v, x) after the op."""
    idx = i64(idx)
    dt = np.float32 if np.asarray(v).dtype == np.float32 else np.float64
    sfx = "f32" if dt == np.float32 else "f64"
    v = np.ascontiguousarray(v, dtype=dt).copy()
    x = np.ascontiguousarray(x, dtype=dt).copy()
    n = idx.shape[0]
    L = lib()
    if op == "dot":
        return dt(getattr(L, f"oracle_sparse_dot_{sfx}")(n, _p(idx), _p(v), _p(x), inc)), v, x
    if op == "axpy":
        getattr(L, f"oracle_sparse_axpy_{sfx}")(n, alpha, _p(idx), _p(v), _p(x), inc)
    else:
        getattr(L, f"oracle_sparse_{op}_{sfx}")(n, _p(idx), _p(v), _p(x), inc)
    return None, v, x


def ref_level1(op, idx, v, x, extent, inc=1, alpha=1.0):
    idx = i64(idx)
    dt = np.float32 if np.asarray(v).dtype == np.float32 else np.float64
    v = np.ascontiguousarray(v, dtype=dt).copy()
    x = np.ascontiguousarray(x, dtype=dt).copy()
    dot = np.zeros(1, dt)
    fn = ref().ref_level1_f32 if dt == np.float32 else ref().ref_level1_f64
    rc = fn(LEVEL1_OPS.index(op), idx.shape[0], extent, _p(idx), _p(v), _p(x), inc, alpha, _p(dot))
    return rc, (dot[0] if op == "dot" else None), v, x


# ---- triangular solve ------------------------------------------------------------
def trsv(fmt, op, lower, unit, n, offsets, indices, values, alpha, b):
    """Oracle triangular solve on a copy of b (n x k row-major, or a vector);
    fmt 0 CSR / 1 CSC, op 0 N / 1 T / 2 H."""
    offsets, indices = i64(offsets), i64(indices)
    dt = np.float32 if values.dtype == np.float32 else np.float64
    values = np.ascontiguousarray(values, dtype=dt)
    out = np.array(b, dtype=dt, copy=True, order="C")
    k = 1 if out.ndim == 1 else out.shape[1]
    fn = lib().oracle_trsv_f32 if dt == np.float32 else lib().oracle_trsv_f64
    fn(fmt, op, 1 if lower else 0, 1 if unit else 0, n, _p(offsets), _p(indices), _p(values), alpha,
       _p(out), k, 1, k)
    return out


def ref_trsv(fmt, op, lower, unit, n, offsets, indices, values, alpha, b):
    offsets, indices = i64(offsets), i64(indices)
    dt = np.float32 if values.dtype == np.float32 else np.float64
    values = np.ascontiguousarray(values, dtype=dt)
    out = np.array(b, dtype=dt, copy=True, order="C")
    k = 1 if out.ndim == 1 else out.shape[1]
    fn = ref().ref_trsv_f32 if dt == np.float32 else ref().ref_trsv_f64
    rc = fn(fmt, op, 1 if lower else 0, 1 if unit else 0, n, indices.shape[0], _p(offsets),
            _p(indices), _p(values), alpha, _p(out), k, 1)
    return rc, out
