"""C ABI checks that need no GPU.

The library loads and exports every function include/asc_sparse_hip.h
declares; a validation-only context (device -1) runs the full rejection
matrix -- the analogue of the reference's CPU-side provider rejection tests
(tests/sparse/blas_test.cc:298-443, sparse_cuda_test.cc:1010-1093) -- and
answers UNAVAILABLE only where work would be enqueued.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import asc_sparse_hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "asc_sparse_hip.h")).read()
    return sorted(set(re.findall(r"\b(asc_sparse_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = H.load()
    names = header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(H.exported_symbols())


def test_device_count_without_gpu_is_zero_or_more():
    assert H.device_count() >= 0


# ---- rejection matrix on a validation-only context ---------------------------------

FAKE = 0x10000  # plausible, never dereferenced: validation-only contexts do not touch memory


def dev_csr(rows=3, cols=4, nnz=5, kind=H.F64, trusted=1, space=H.DEVICE, fmt=H.CSR,
            index_bytes=8, base=FAKE):
    return H.SparseDesc(base, base + 0x1000, base + 0x2000, rows, cols, nnz, fmt, kind,
                        index_bytes, 8, space, trusted)


def dev_vec(n, kind=H.F64, base=0x100000, stride=1, writable=1):
    return H.VectorDesc(base, n, stride, kind, H.DEVICE, writable)


def spmv_code(vctx, a, x, y, alg=H.ALG_AUTO, acc=H.ACC_PLAIN, ws=None, ws_bytes=0):
    opt = H.options(alg, acc)
    return vctx.lib.asc_sparse_hip_csr_spmv(vctx.handle, 1.0, C.byref(a), C.byref(x), 0.0,
                                            C.byref(y), ws, ws_bytes, C.byref(opt)).code


def test_valid_call_reaches_enqueue(vctx):
    assert spmv_code(vctx, dev_csr(), dev_vec(4), dev_vec(3, base=0x200000),
                     alg=H.ALG_VECTOR[8]) == H.UNAVAILABLE


def test_rejects_untrusted_structure(vctx):
    assert spmv_code(vctx, dev_csr(trusted=0), dev_vec(4), dev_vec(3, base=0x200000)) == H.INVALID_ARGUMENT


def test_rejects_host_matrix(vctx):
    assert spmv_code(vctx, dev_csr(space=H.HOST), dev_vec(4), dev_vec(3, base=0x200000)) == H.MEMORY_ACCESS


def test_rejects_shape_mismatch(vctx):
    assert spmv_code(vctx, dev_csr(), dev_vec(5), dev_vec(3, base=0x200000)) == H.SHAPE
    assert spmv_code(vctx, dev_csr(), dev_vec(4), dev_vec(2, base=0x200000)) == H.SHAPE


def test_rejects_nnz_beyond_shape(vctx):
    assert spmv_code(vctx, dev_csr(nnz=13), dev_vec(4), dev_vec(3, base=0x200000)) == H.SHAPE


def test_rejects_output_overlap(vctx):
    a = dev_csr()
    # y placed over the values array
    assert spmv_code(vctx, a, dev_vec(4), dev_vec(3, base=FAKE + 0x2000)) == H.INVALID_ARGUMENT
    # y overlapping x
    assert spmv_code(vctx, a, dev_vec(4), dev_vec(3, base=0x100008)) == H.INVALID_ARGUMENT


def test_rejects_wrong_vector_kind_and_csc(vctx):
    assert spmv_code(vctx, dev_csr(), dev_vec(4, kind=H.F32), dev_vec(3, base=0x200000)) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, dev_csr(fmt=H.CSC), dev_vec(4), dev_vec(3, base=0x200000)) == H.INVALID_ARGUMENT


def test_rejects_readonly_output_and_zero_stride(vctx):
    assert spmv_code(vctx, dev_csr(), dev_vec(4), dev_vec(3, base=0x200000, writable=0)) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, dev_csr(), dev_vec(4, stride=0), dev_vec(3, base=0x200000)) == H.INVALID_ARGUMENT


def test_rejects_unknown_algorithm_and_bad_modes(vctx):
    a, x, y = dev_csr(), dev_vec(4), dev_vec(3, base=0x200000)
    assert spmv_code(vctx, a, x, y, alg=42) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, a, x, y, acc=H.ACC_TWOSUM) == H.UNSUPPORTED  # FP64 storage
    b = dev_csr(kind=H.BF16)
    xf, yf = dev_vec(4, kind=H.F32), dev_vec(3, kind=H.F32, base=0x200000)
    assert spmv_code(vctx, b, xf, yf, alg=H.ALG_ORDERED) == H.UNSUPPORTED
    assert spmv_code(vctx, b, xf, yf, alg=H.ALG_VECTOR[4], acc=H.ACC_TWOSUM) == H.UNAVAILABLE


def test_merge_requires_workspace(vctx):
    a, x, y = dev_csr(), dev_vec(4), dev_vec(3, base=0x200000)
    need = C.c_size_t()
    opt = H.options(H.ALG_MERGE)
    assert vctx.lib.asc_sparse_hip_csr_spmv_workspace_size(
        vctx.handle, C.byref(a), C.byref(x), C.byref(y), C.byref(opt), C.byref(need)).code == H.OK
    assert need.value > 0
    assert spmv_code(vctx, a, x, y, alg=H.ALG_MERGE) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, a, x, y, alg=H.ALG_MERGE, ws=0x400000, ws_bytes=need.value - 1) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, a, x, y, alg=H.ALG_MERGE, ws=0x400000 + 8, ws_bytes=need.value) == H.INVALID_ARGUMENT
    assert spmv_code(vctx, a, x, y, alg=H.ALG_MERGE, ws=0x400000, ws_bytes=need.value) == H.UNAVAILABLE
    # workspace overlapping the output
    assert spmv_code(vctx, a, x, y, alg=H.ALG_MERGE, ws=0x200000, ws_bytes=need.value) == H.INVALID_ARGUMENT


def test_int32_indices_cannot_address_large_extent(vctx):
    a = dev_csr(rows=3, cols=(1 << 31) + 5, nnz=5, index_bytes=4)
    assert spmv_code(vctx, a, dev_vec((1 << 31) + 5), dev_vec(3, base=0x200000)) == H.OVERFLOW


def test_null_and_dead_contexts():
    lib = H.load()
    a, x, y = dev_csr(), dev_vec(4), dev_vec(3, base=0x200000)
    st = lib.asc_sparse_hip_csr_spmv(None, 1.0, C.byref(a), C.byref(x), 0.0, C.byref(y), None, 0, None)
    assert st.code == H.INVALID_ARGUMENT
    c = H.Context(-1)
    c.close()


def test_standard_spmv_validation(vctx):
    a = dev_csr()
    x, y = dev_vec(3), dev_vec(4, base=0x200000)  # op T: x has rows, y has cols
    opt = H.options(H.ALG_ORDERED)
    need = C.c_size_t()
    st = vctx.lib.asc_sparse_hip_standard_spmv_workspace_size(
        vctx.handle, H.OP_T, C.byref(a), C.byref(x), C.byref(y), C.byref(opt), C.byref(need))
    assert st.code == H.OK and need.value > 0  # CSR with op T transposes in the workspace
    st = vctx.lib.asc_sparse_hip_standard_spmv(vctx.handle, 7, H.Scalar(1, 0), C.byref(a),
                                               C.byref(x), C.byref(y), None, 0, C.byref(opt))
    assert st.code == H.INVALID_ARGUMENT
    st = vctx.lib.asc_sparse_hip_standard_spmv(vctx.handle, H.OP_N, H.Scalar(1, 0), C.byref(a),
                                               C.byref(x), C.byref(y), None, 0, C.byref(opt))
    assert st.code == H.SHAPE


def test_clone_validates_host_structure(vctx):
    off = np.array([0, 2, 4, 5], np.int64)
    bad = np.array([3, 1, 0, 2, 1], np.int64)  # first row unsorted
    good = np.array([1, 3, 0, 2, 1], np.int64)
    vals = np.ones(5)
    out = H.SparseDesc()
    for idx, code in ((bad, H.INVALID_ARGUMENT), (good, H.UNAVAILABLE)):
        src = H.host_sparse(H.CSR, 3, 4, off, idx, vals, H.F64)
        st = vctx.lib.asc_sparse_hip_clone_compressed(vctx.handle, C.byref(src), 4, FAKE,
                                                      FAKE + 0x100, FAKE + 0x200, C.byref(out))
        assert st.code == code
    oob = np.array([1, 4, 0, 2, 1], np.int64)
    src = H.host_sparse(H.CSR, 3, 4, off, oob, vals, H.F64)
    st = vctx.lib.asc_sparse_hip_clone_compressed(vctx.handle, C.byref(src), 8, FAKE,
                                                  FAKE + 0x100, FAKE + 0x200, C.byref(out))
    assert st.code == H.INDEX
    # overlapping destinations
    src = H.host_sparse(H.CSR, 3, 4, off, good, vals, H.F64)
    st = vctx.lib.asc_sparse_hip_clone_compressed(vctx.handle, C.byref(src), 8, FAKE,
                                                  FAKE + 8, FAKE + 0x200, C.byref(out))
    assert st.code == H.INVALID_ARGUMENT


def test_conversion_validation(vctx):
    coo = H.SparseDesc(FAKE, None, FAKE + 0x1000, 3, 4, 5, H.COO, H.F64, 8, 8, H.DEVICE, 1)
    need = C.c_size_t()
    assert vctx.lib.asc_sparse_hip_convert_coordinate_workspace_size(
        vctx.handle, C.byref(coo), H.CSC, 8, C.byref(need)).code == H.OK
    assert vctx.lib.asc_sparse_hip_convert_coordinate_workspace_size(
        vctx.handle, C.byref(coo), 9, 8, C.byref(need)).code == H.INVALID_ARGUMENT
    csr = dev_csr()
    assert vctx.lib.asc_sparse_hip_convert_compressed_workspace_size(
        vctx.handle, C.byref(csr), C.byref(need)).code == H.OK
    untrusted = dev_csr(trusted=0)
    assert vctx.lib.asc_sparse_hip_convert_compressed_workspace_size(
        vctx.handle, C.byref(untrusted), C.byref(need)).code == H.INVALID_ARGUMENT


def test_fault_injection_arms_and_disarms(vctx):
    H.test_inject(H.INJECT_ALLOCATION, 0)
    with pytest.raises(H.AscHipError) as e:
        vctx.alloc(64, H.HOST)
    assert e.value.code == H.ALLOCATION
    buf = vctx.alloc(64, H.HOST)  # disarmed after firing
    buf.free()
    H.test_inject(H.INJECT_ALLOCATION, 1)
    b1 = vctx.alloc(64, H.HOST)
    with pytest.raises(H.AscHipError):
        vctx.alloc(64, H.HOST)
    b1.free()
    H.test_inject(H.INJECT_ALLOCATION, -1)


def test_panel_analysis_validation(vctx):
    lib, h = vctx.lib, vctx.handle
    need = C.c_size_t()

    def size(a, opts=None):
        st = lib.asc_sparse_hip_csr_spmv_analysis_size(h, C.byref(a), C.byref(opts) if opts else None,
                                                       C.byref(need))
        return st.code

    def aligned(n):
        return (n + 255) // 256 * 256

    def scratch(rows):  # flags + positions + count + scan partials (tiles of 4096)
        return 2 * aligned(4 * rows) + 256 + aligned(4 * max(1, (rows + 4095) // 4096))

    def layout(rows, panels):  # row list, row starts, split table, scratch
        return aligned(4 * rows) + aligned(8 * rows) + aligned(panels * rows * 4) + scratch(rows)

    big = dev_csr(rows=1000, cols=10_000_000, nnz=5000)
    # default width 2^18 FP64 columns -> 39 panels, rounded up to 40 (one per XCD at a time)
    assert size(big) == H.OK
    assert need.value == layout(1000, 40)
    f32 = dev_csr(rows=1000, cols=10_000_000, nnz=5000, kind=H.F32)
    assert size(f32) == H.OK  # 2^19 FP32 columns -> 20 panels -> 24
    assert need.value == layout(1000, 24)
    assert size(big, H.panel_options(4_000_000)) == H.OK  # 3 panels -> 8
    assert need.value == layout(1000, 8)
    # LDS staging: a panel must fit in 160 KiB of x (20480 FP64 / 40960 FP32 columns)
    # (the width is rounded to a multiple of 8 panels: 20480 -> 496 panels of 20162 columns)
    assert size(big, H.panel_options(20480, staging=H.STAGING_LDS)) == H.OK
    assert need.value == layout(1000, 496)
    assert size(big, H.panel_options(30000, staging=H.STAGING_LDS)) == H.INVALID_ARGUMENT
    assert size(big, H.panel_options(0, staging=H.STAGING_LDS)) == H.OK  # LDS-sized default
    assert need.value == layout(1000, 496)
    # AUTO takes LDS staging once a row averages 48 entries per LDS-sized panel
    dense_rows = dev_csr(rows=1000, cols=10_000_000, nnz=1000 * 48 * 496)
    assert size(dense_rows) == H.OK and need.value == layout(1000, 496)
    sparser = dev_csr(rows=1000, cols=10_000_000, nnz=1000 * 47 * 496)
    assert size(sparser) == H.OK and need.value == layout(1000, 40)
    assert size(f32, H.panel_options(40960, staging=H.STAGING_LDS)) == H.OK
    assert size(f32, H.panel_options(50000, staging=H.STAGING_LDS)) == H.INVALID_ARGUMENT
    assert size(big, H.panel_options(0, staging=4)) == H.INVALID_ARGUMENT
    # stream staging: panels of at most 128 KiB of x and 32768 columns; without a
    # device every row counts as listed (the bound), and entries are numbered
    # in 32 bits
    assert size(big, H.panel_options(0, staging=H.STAGING_STREAM)) == H.OK
    stream_bound = need.value
    assert stream_bound > (5000 + 616 * 512) * 10  # the padded stream copy at least
    assert size(big, H.panel_options(16384, staging=H.STAGING_STREAM)) == H.OK
    assert size(big, H.panel_options(16385, staging=H.STAGING_STREAM)) == H.INVALID_ARGUMENT
    assert size(f32, H.panel_options(32768, staging=H.STAGING_STREAM)) == H.OK
    assert size(f32, H.panel_options(32769, staging=H.STAGING_STREAM)) == H.INVALID_ARGUMENT
    many = dev_csr(rows=5_000_000, cols=10_000_000, nnz=1 << 32)  # entries past 32 bits
    assert size(many, H.panel_options(0, staging=H.STAGING_STREAM)) == H.UNSUPPORTED
    assert size(dev_csr()) == H.OK  # 4 columns: still 8 (tiny) panels
    assert size(big, H.panel_options(-1)) == H.INVALID_ARGUMENT
    assert size(big, H.panel_options(0, -5)) == H.INVALID_ARGUMENT
    assert size(big, H.panel_options(1)) == H.INVALID_ARGUMENT  # > 1024 panels
    assert size(dev_csr(fmt=H.CSC)) == H.INVALID_ARGUMENT
    assert size(dev_csr(trusted=0)) == H.INVALID_ARGUMENT
    assert size(dev_csr(kind=H.C128)) == H.UNSUPPORTED
    ws = 0x4000000
    assert size(big) == H.OK
    n = need.value

    def analyze(opts, ptr, nbytes):
        return lib.asc_sparse_hip_csr_spmv_analyze(h, C.byref(big), C.byref(opts) if opts else None,
                                                   ptr, nbytes).code

    # bad kernels, missing/short/overlapping buffer, then the enqueue point
    assert analyze(H.panel_options(0, 0, H.ALG_MERGE), ws, n) == H.INVALID_ARGUMENT
    assert analyze(H.panel_options(0, 0, H.ALG_AUTO, H.ALG_PANEL), ws, n) == H.INVALID_ARGUMENT
    assert analyze(None, None, 0) == H.INVALID_ARGUMENT
    assert analyze(None, ws, n - 4) == H.INVALID_ARGUMENT
    assert analyze(None, FAKE, n) == H.INVALID_ARGUMENT
    assert analyze(H.panel_options(0, 16, H.ALG_VECTOR[2], H.ALG_VECTOR[32]), ws, n) == H.UNAVAILABLE
    assert analyze(H.panel_options(0, 16, staging=H.STAGING_STREAM), ws, stream_bound) == H.UNAVAILABLE


def test_panel_algorithm_requires_analysis(vctx):
    a = dev_csr()
    # no analysis registered: explicit ALG_PANEL is rejected before enqueue
    assert spmv_code(vctx, a, dev_vec(4), dev_vec(3, base=0x200000), alg=H.ALG_PANEL) == H.INVALID_ARGUMENT
    # plan_set cannot register a panel plan without its split table
    assert vctx.lib.asc_sparse_hip_plan_set(vctx.handle, C.byref(a), H.ALG_PANEL).code == H.INVALID_ARGUMENT
    assert vctx.lib.asc_sparse_hip_plan_set(vctx.handle, C.byref(a), 11).code == H.INVALID_ARGUMENT


def test_finalize_validation(vctx):
    lib, h = vctx.lib, vctx.handle
    need = C.c_size_t()

    def trip(rows=3, cols=4, nnz=5, kind=H.F64, space=H.DEVICE, fmt=H.COO, base=FAKE):
        return H.SparseDesc(base, None, base + 0x1000, rows, cols, nnz, fmt, kind, 8, 8, space, 0)

    def size(t):
        return lib.asc_sparse_hip_coo_finalize_workspace_size(h, C.byref(t), C.byref(need)).code

    assert size(trip()) == H.OK  # untrusted raw triplets are the input
    assert size(trip(nnz=50)) == H.OK  # duplicates may exceed rows*cols
    assert size(trip(kind=H.BF16)) == H.UNSUPPORTED
    assert size(trip(fmt=H.CSR)) == H.INVALID_ARGUMENT
    assert size(trip(space=H.HOST)) == H.MEMORY_ACCESS
    assert size(trip(rows=-1)) == H.SHAPE
    assert size(trip(cols=1 << 33)) == H.UNSUPPORTED
    assert size(trip()) == H.OK
    n = need.value
    out = H.SparseDesc()
    ws, oc, ov = 0x4000000, 0x8000000, 0x9000000

    def fin(t, dup, zero, c=oc, v=ov, w=ws, nb=None):
        return lib.asc_sparse_hip_coo_finalize(h, C.byref(t), dup, zero, c, v, C.byref(out), w,
                                               n if nb is None else nb).code

    assert fin(trip(), 2, 0) == H.INVALID_ARGUMENT  # unknown duplicate policy
    assert fin(trip(), 1, 2) == H.INVALID_ARGUMENT  # unknown zero policy
    assert fin(trip(), 1, 0, c=None) == H.INVALID_ARGUMENT
    assert fin(trip(), 1, 0, c=FAKE) == H.INVALID_ARGUMENT  # output over the triplets
    assert fin(trip(), 1, 0, v=oc + 8) == H.INVALID_ARGUMENT  # outputs overlap
    assert fin(trip(), 1, 0, nb=n - 8) == H.INVALID_ARGUMENT  # short workspace
    assert fin(trip(), 1, 1) == H.UNAVAILABLE  # valid: reaches the enqueue point


def test_level1_validation(vctx):
    lib, h = vctx.lib, vctx.handle

    def ix(nnz=2, extent=3, kind=H.F64, space=H.DEVICE, trusted=1, writable=1, base=FAKE):
        return H.IndexedDesc(base, base + 0x1000, nnz, extent, kind, space, writable, trusted)

    def call(op, sp, dense, result=None, alg=H.ALG_AUTO, conj=0):
        opt = H.options(alg)
        return lib.asc_sparse_hip_standard_level1(h, op, conj, H.Scalar(1.0, 0.0), C.byref(sp),
                                                  C.byref(dense),
                                                  C.byref(result) if result is not None else None,
                                                  C.byref(opt)).code

    d3 = dev_vec(3, base=0x200000)
    r1 = dev_vec(1, base=0x300000)
    assert call(H.LEVEL1_DOT, ix(), d3, r1) == H.UNAVAILABLE
    assert call(H.LEVEL1_DOT, ix(), d3, r1, alg=H.ALG_ORDERED) == H.UNAVAILABLE
    assert call(H.LEVEL1_DOT, ix(), d3, r1, alg=H.ALG_MERGE) == H.INVALID_ARGUMENT
    assert call(H.LEVEL1_DOT, ix(), d3, None) == H.INVALID_ARGUMENT  # no result vector
    assert call(H.LEVEL1_DOT, ix(), d3, dev_vec(2, base=0x300000)) == H.SHAPE
    assert call(H.LEVEL1_DOT, ix(), d3, dev_vec(1, base=0x200000)) == H.INVALID_ARGUMENT  # overlap
    assert call(5, ix(), d3) == H.INVALID_ARGUMENT
    assert call(H.LEVEL1_AXPY, ix(), d3, conj=2) == H.INVALID_ARGUMENT
    assert call(H.LEVEL1_AXPY, ix(trusted=0), d3) == H.INVALID_ARGUMENT
    assert call(H.LEVEL1_AXPY, ix(space=H.HOST), d3) == H.MEMORY_ACCESS
    assert call(H.LEVEL1_AXPY, ix(kind=H.C128), d3) == H.UNSUPPORTED
    assert call(H.LEVEL1_AXPY, ix(nnz=4), d3) == H.INVALID_ARGUMENT  # nnz > dense extent
    assert call(H.LEVEL1_AXPY, ix(), dev_vec(4, base=0x200000)) == H.SHAPE
    assert call(H.LEVEL1_AXPY, ix(), dev_vec(3, base=0x200000, writable=0)) == H.INVALID_ARGUMENT
    assert call(H.LEVEL1_AXPY, ix(), dev_vec(3, base=FAKE + 0x1000)) == H.INVALID_ARGUMENT  # overlap
    assert call(H.LEVEL1_GATHER, ix(writable=0), d3) == H.INVALID_ARGUMENT
    for op in (H.LEVEL1_AXPY, H.LEVEL1_GATHER, H.LEVEL1_GATHER_ZERO, H.LEVEL1_SCATTER):
        assert call(op, ix(), d3) == H.UNAVAILABLE
    # host clone validation: unsorted / out-of-range indices
    idx = np.array([0, 2, 1], np.int64)
    vals = np.zeros(3)
    out = H.IndexedDesc()
    src = H.IndexedDesc(idx.ctypes.data, vals.ctypes.data, 3, 5, H.F64, H.HOST, 0, 0)
    assert lib.asc_sparse_hip_clone_indexed(h, C.byref(src), FAKE, FAKE + 0x100, C.byref(out)).code == H.INDEX
    idx2 = np.array([0, 2, 5], np.int64)
    src = H.IndexedDesc(idx2.ctypes.data, vals.ctypes.data, 3, 5, H.F64, H.HOST, 0, 0)
    assert lib.asc_sparse_hip_clone_indexed(h, C.byref(src), FAKE, FAKE + 0x100, C.byref(out)).code == H.INDEX
    idx3 = np.array([0, 2, 4], np.int64)
    src = H.IndexedDesc(idx3.ctypes.data, vals.ctypes.data, 3, 5, H.F64, H.HOST, 0, 0)
    assert lib.asc_sparse_hip_clone_indexed(h, C.byref(src), FAKE, FAKE + 0x100, C.byref(out)).code == H.UNAVAILABLE


def test_triangular_validation(vctx):
    lib, h = vctx.lib, vctx.handle
    need = C.c_size_t()
    sq = dev_csr(rows=4, cols=4, nnz=6)
    assert lib.asc_sparse_hip_triangular_analysis_size(h, H.OP_N, C.byref(sq), C.byref(need)).code == H.OK
    n_direct = need.value
    assert lib.asc_sparse_hip_triangular_analysis_size(h, H.OP_T, C.byref(sq), C.byref(need)).code == H.OK
    assert need.value > n_direct  # op T on CSR keeps a transposed copy
    assert lib.asc_sparse_hip_triangular_analysis_size(h, 3, C.byref(sq), C.byref(need)).code == H.INVALID_ARGUMENT
    assert lib.asc_sparse_hip_triangular_analysis_size(h, H.OP_N, C.byref(dev_csr()), C.byref(need)).code == H.SHAPE
    assert lib.asc_sparse_hip_triangular_analysis_size(
        h, H.OP_N, C.byref(dev_csr(rows=4, cols=4, nnz=6, kind=H.BF16)), C.byref(need)).code == H.UNSUPPORTED
    info = H.TriangularInfo()
    ws = 0x4000000
    lib.asc_sparse_hip_triangular_analysis_size(h, H.OP_N, C.byref(sq), C.byref(need))
    assert lib.asc_sparse_hip_triangular_analyze(h, H.OP_N, 2, 0, C.byref(sq), ws, need.value,
                                                 C.byref(info)).code == H.INVALID_ARGUMENT
    assert lib.asc_sparse_hip_triangular_analyze(h, H.OP_N, 1, 2, C.byref(sq), ws, need.value,
                                                 C.byref(info)).code == H.INVALID_ARGUMENT
    assert lib.asc_sparse_hip_triangular_analyze(h, H.OP_N, 1, 0, C.byref(sq), ws, need.value - 8,
                                                 C.byref(info)).code == H.INVALID_ARGUMENT
    assert lib.asc_sparse_hip_triangular_analyze(h, H.OP_N, 1, 0, C.byref(sq), ws, need.value,
                                                 C.byref(info)).code == H.UNAVAILABLE
    # a solve with an info that does not describe this matrix
    rhs = H.MatrixDesc(0x6000000, 4, 1, 1, H.ROW_MAJOR, H.F64, H.DEVICE, 1)
    bogus = H.TriangularInfo(4, 6, 1, 4)
    assert lib.asc_sparse_hip_triangular_solve(h, H.OP_N, H.Scalar(1.0, 0.0), C.byref(sq), C.byref(rhs),
                                               ws, need.value, C.byref(bogus)).code == H.INVALID_ARGUMENT
    good = H.TriangularInfo(4, 6, 1, 4, sq.outer_offsets, sq.inner_indices, sq.values, H.OP_N, 1, 0, H.CSR)
    assert lib.asc_sparse_hip_triangular_solve(h, H.OP_N, H.Scalar(1.0, 0.0), C.byref(sq), C.byref(rhs),
                                               ws, need.value, C.byref(good)).code == H.UNAVAILABLE
    wrong_rows = H.MatrixDesc(0x6000000, 3, 1, 1, H.ROW_MAJOR, H.F64, H.DEVICE, 1)
    assert lib.asc_sparse_hip_triangular_solve(h, H.OP_N, H.Scalar(1.0, 0.0), C.byref(sq),
                                               C.byref(wrong_rows), ws, need.value,
                                               C.byref(good)).code == H.SHAPE


def test_span_byte_counts_that_overflow_are_refused(vctx):
    """Array, strided-vector and dense-block byte counts must be representable
    (CheckedByteCount, compressed.h:477-497): an operand whose span would wrap
    is refused before its allocation is looked at."""
    big = 1 << 61
    # (rows + 1) int64 offsets overflow: rows = 2^61, one column, no entries
    a = dev_csr(rows=big, cols=1, nnz=0)
    x, y = dev_vec(1), dev_vec(big)
    assert spmv_code(vctx, a, x, y, alg=H.ALG_ORDERED) == H.OVERFLOW
    # a stride whose span wraps
    a = dev_csr()
    assert spmv_code(vctx, a, dev_vec(4, stride=1 << 62), dev_vec(3), alg=H.ALG_ORDERED) == H.OVERFLOW
    assert spmv_code(vctx, a, dev_vec(4), dev_vec(3, stride=-(1 << 62)), alg=H.ALG_ORDERED) == H.OVERFLOW
    assert spmv_code(vctx, a, dev_vec(4), dev_vec(3, stride=-(2 ** 63)), alg=H.ALG_ORDERED) == H.OVERFLOW
    # large strides that still fit pass validation (a validation-only context
    # then refuses to enqueue)
    assert spmv_code(vctx, a, dev_vec(4, stride=1 << 40), dev_vec(3, stride=-(1 << 40)),
                     alg=H.ALG_ORDERED) == H.UNAVAILABLE
    # a dense block whose leading dimension makes the span wrap
    b = H.MatrixDesc(0x200000, 4, 2, 1 << 62, H.COLUMN_MAJOR, H.F64, H.DEVICE, 0)
    c = H.MatrixDesc(0x300000, 3, 2, 3, H.COLUMN_MAJOR, H.F64, H.DEVICE, 1)
    st = vctx.lib.asc_sparse_hip_standard_spmm(vctx.handle, H.OP_N, H.Scalar(1.0, 0.0), C.byref(a),
                                               C.byref(b), C.byref(c), None, 0, None)
    assert st.code == H.OVERFLOW


def test_spmm_form_options(vctx):
    """SpMM takes ALG_AUTO (plain) or ALG_ORDERED with plain accumulation;
    other algorithms and compensated accumulation are refused before any
    operand is looked at; NULL options are AUTO."""
    a = dev_csr()
    b = H.MatrixDesc(0x200000, 4, 2, 2, H.ROW_MAJOR, H.F64, H.DEVICE, 0)
    c = H.MatrixDesc(0x300000, 3, 2, 2, H.ROW_MAJOR, H.F64, H.DEVICE, 1)
    need = C.c_size_t()

    def size_code(opt):
        return vctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
            vctx.handle, H.OP_N, C.byref(a), C.byref(b), C.byref(c),
            None if opt is None else C.byref(opt), C.byref(need)).code

    def call_code(opt):
        return vctx.lib.asc_sparse_hip_standard_spmm(
            vctx.handle, H.OP_N, H.Scalar(1.0, 0.0), C.byref(a), C.byref(b), C.byref(c), None, 0,
            None if opt is None else C.byref(opt)).code

    for opt in (None, H.options(H.ALG_AUTO), H.options(H.ALG_ORDERED)):
        assert size_code(opt) == H.OK
        assert call_code(opt) == H.UNAVAILABLE  # validation-only context
    for alg in (H.ALG_MERGE, H.ALG_VECTOR[8], 200):
        assert size_code(H.options(alg)) == H.INVALID_ARGUMENT
        assert call_code(H.options(alg)) == H.INVALID_ARGUMENT
    assert size_code(H.options(H.ALG_AUTO, H.ACC_TWOSUM)) == H.UNSUPPORTED
    assert call_code(H.options(H.ALG_ORDERED, H.ACC_TWOSUM)) == H.UNSUPPORTED
    # ALG_PANEL (the plain form over a stream-staged column-panel analysis):
    # refused for operands the plain form does not take (k = 2 here), for
    # op(A) = A^T, and without an analysis of the matrix.
    assert size_code(H.options(H.ALG_PANEL)) == H.UNSUPPORTED
    b4 = H.MatrixDesc(0x200000, 4, 4, 4, H.ROW_MAJOR, H.F64, H.DEVICE, 0)
    c4 = H.MatrixDesc(0x300000, 3, 4, 4, H.ROW_MAJOR, H.F64, H.DEVICE, 1)
    for op, code in ((H.OP_N, H.INVALID_ARGUMENT), (H.OP_T, H.UNSUPPORTED)):
        cc = c4 if op == H.OP_N else H.MatrixDesc(0x300000, 4, 4, 4, H.ROW_MAJOR, H.F64, H.DEVICE, 1)
        bb = b4 if op == H.OP_N else H.MatrixDesc(0x200000, 3, 4, 4, H.ROW_MAJOR, H.F64, H.DEVICE, 0)
        st = vctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
            vctx.handle, op, C.byref(a), C.byref(bb), C.byref(cc),
            C.byref(H.options(H.ALG_PANEL)), C.byref(need))
        assert st.code == code, (op, st.code)
