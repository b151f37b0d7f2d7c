"""GPU parity of the standardized SpMV (y += alpha*op(A)*x) and SpMM
(C += alpha*op(A)*B) through the C ABI, CSR and CSC, op N/T/H.

Ordered SpMV and every SpMM are bit-exact with the oracle; plain SpMV stays
within 4*len*eps*|alpha|*(|A||x|)_i.  Golden vectors: standard_blas_test.cc.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
KIND = {np.float64: H.F64, np.float32: H.F32}
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}


def device_matrix(ctx, fmt, rows, cols, off, idx, val, index_bytes=8):
    kind = KIND[val.dtype.type]
    return H.clone_compressed(ctx, fmt, rows, cols, off, idx, val, kind, index_bytes)


def run_spmv(ctx, op, alpha, a, x, y0, alg):
    kind = a.desc.element_kind
    dt = np.float64 if kind == H.F64 else np.float32
    xb, yb = ctx.upload(x.astype(dt)), ctx.upload(y0.astype(dt))
    H.standard_spmv(ctx, op, alpha, a, H.vector(xb, len(x), kind, writable=False),
                    H.vector(yb, len(y0), kind), alg=alg)
    return ctx.download(yb, dt, len(y0))


def run_spmm(ctx, op, alpha, a, b, c0, layout):
    kind = a.desc.element_kind
    dt = np.float64 if kind == H.F64 else np.float32
    nb, k = b.shape
    nc = c0.shape[0]
    if layout == H.ROW_MAJOR:
        bb, cc, ldb, ldc = b.reshape(-1), c0.reshape(-1), k, k
    else:
        bb, cc, ldb, ldc = b.T.reshape(-1), c0.T.reshape(-1), nb, nc
    bbuf, cbuf = ctx.upload(bb.astype(dt)), ctx.upload(cc.astype(dt))
    H.standard_spmm(ctx, op, alpha, a, H.dense(bbuf, nb, k, ldb, layout, kind, writable=False),
                    H.dense(cbuf, nc, k, ldc, layout, kind))
    out = ctx.download(cbuf, dt, nc * k)
    return (out.reshape(nc, k) if layout == H.ROW_MAJOR else out.reshape(k, nc).T), (bb, cc, ldb, ldc)


def test_standard_golden(ctx):
    g = GOLD["standard"]
    rows, cols = g["shape"]
    vals = np.array(g["values"], dtype=np.float64)
    x = np.array(g["x"], dtype=np.float64)
    for fmt, off, idx in ((H.CSR, g["csr_offsets"], g["csr_indices"]), (H.CSC, g["csc_offsets"], g["csc_indices"])):
        a = device_matrix(ctx, fmt, rows, cols, np.array(off), np.array(idx), vals)
        for alg in (H.ALG_ORDERED, H.ALG_MERGE, H.ALG_VECTOR[4]):
            np.testing.assert_array_equal(run_spmv(ctx, H.OP_N, 1.0, a, x, np.zeros(2), alg), g["spmv_n"])
            np.testing.assert_array_equal(run_spmv(ctx, H.OP_H, 1.0, a, x, np.zeros(2), alg), g["spmv_t"])
    nanm = device_matrix(ctx, H.CSR, rows, cols, np.array(g["csr_offsets"]), np.array(g["csr_indices"]),
                         np.full(3, np.nan))
    np.testing.assert_array_equal(run_spmv(ctx, H.OP_N, 0.0, nanm, x, np.array([7.0, 9.0]), H.ALG_MERGE), [7, 9])
    a = device_matrix(ctx, H.CSR, rows, cols, np.array(g["csr_offsets"]), np.array(g["csr_indices"]), vals)
    c, _ = run_spmm(ctx, H.OP_N, 1.0, a, np.array(g["spmm_row_major_b"], float).reshape(2, 2),
                    np.zeros((2, 2)), H.ROW_MAJOR)
    np.testing.assert_array_equal(c.reshape(-1), g["spmm_row_major_c"])
    csc = device_matrix(ctx, H.CSC, rows, cols, np.array(g["csc_offsets"]), np.array(g["csc_indices"]), vals)
    bcm = np.array(g["spmm_csc_h_col_major_b"], float).reshape(2, 2).T  # column-major storage
    c, _ = run_spmm(ctx, H.OP_H, 1.0, csc, bcm, np.zeros((2, 2)), H.COLUMN_MAJOR)
    np.testing.assert_array_equal(c.T.reshape(-1), g["spmm_csc_h_col_major_c"])
    c0 = np.array([5.0, 6.0, 7.0, 8.0]).reshape(2, 2)
    c, _ = run_spmm(ctx, H.OP_N, 0.0, nanm, np.ones((2, 2)), c0, H.ROW_MAJOR)
    np.testing.assert_array_equal(c, c0)


@pytest.mark.parametrize("fmt", [H.CSR, H.CSC])
@pytest.mark.parametrize("op", [H.OP_N, H.OP_T, H.OP_H])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_standard_spmv_random(ctx, fmt, op, dtype):
    m = G.make(G.ZIPF, 6000, 4500, seed=31, values=G.VALUES_NORMAL, zipf_max=3000)
    off, idx, val = m.offsets, m.indices.astype(np.int64), m.values.astype(dtype)
    rows, cols = m.rows, m.cols
    if fmt == H.CSC:
        off, idx, val = O.transpose(0, rows, cols, off, idx, val)
    nin, nout = (cols, rows) if op == H.OP_N else (rows, cols)
    x = G.vector(nin, 5).astype(dtype)
    y0 = G.vector(nout, 6).astype(dtype)
    yref = O.std_spmv(0 if fmt == H.CSR else 1, op, rows, cols, off, idx, val, -0.75, x, y0.copy())
    # row view of op(A) for the plain-mode bound
    if (fmt == H.CSR) == (op == H.OP_N):
        ro, ri, rv = (off, idx, val)
    else:
        ro, ri, rv = O.transpose(0 if fmt == H.CSR else 1, rows, cols, off, idx, val)
    absax = O.abs_spmv(ro, ri, rv.astype(np.float64), x.astype(np.float64))
    lens = np.diff(ro)
    for ib in (4, 8):
        a = device_matrix(ctx, fmt, rows, cols, off, idx, val, ib)
        for alg in (H.ALG_ORDERED, H.ALG_MERGE, H.ALG_VECTOR[1], H.ALG_VECTOR[32]):
            print(f"fmt {fmt} op {op} {dtype.__name__} ib {ib} alg {H.ALG_NAMES[alg]}", flush=True)
            y = run_spmv(ctx, op, -0.75, a, x, y0, alg)
            if alg == H.ALG_ORDERED:
                np.testing.assert_array_equal(y.view(np.uint8), yref.view(np.uint8))
            else:
                bound = 4 * np.maximum(lens, 1) * EPS[dtype] * 0.75 * absax + 4 * EPS[dtype] * np.abs(y0)
                assert np.all(np.abs(y.astype(float) - yref.astype(float)) <= bound)


@pytest.mark.parametrize("fmt", [H.CSR, H.CSC])
@pytest.mark.parametrize("op", [H.OP_N, H.OP_T])
@pytest.mark.parametrize("layout", [H.ROW_MAJOR, H.COLUMN_MAJOR])
@pytest.mark.parametrize("k", [1, 5, 32, 70])
def test_standard_spmm_bit_exact(ctx, fmt, op, layout, k):
    m = G.make(G.ZIPF, 3000, 2500, seed=41, values=G.VALUES_NORMAL, zipf_max=800)
    off, idx, val = m.offsets, m.indices.astype(np.int64), m.values
    rows, cols = m.rows, m.cols
    if fmt == H.CSC:
        off, idx, val = O.transpose(0, rows, cols, off, idx, val)
    nin, nout = (cols, rows) if op == H.OP_N else (rows, cols)
    rng = np.random.default_rng(k)
    b = rng.uniform(-1, 1, (nin, k))
    c0 = rng.uniform(-1, 1, (nout, k))
    a = device_matrix(ctx, fmt, rows, cols, off, idx, val, 4)
    print(f"spmm fmt {fmt} op {op} layout {layout} k {k}", flush=True)
    got, (bb, cc, ldb, ldc) = run_spmm(ctx, op, 1.5, a, b, c0, layout)
    ref = O.std_spmm(0 if fmt == H.CSR else 1, op, rows, cols, off, idx, val, 1.5, bb.copy(), layout,
                     ldb, k, cc.copy(), layout, ldc)
    ref = ref.reshape(nout, k) if layout == H.ROW_MAJOR else ref.reshape(k, nout).T
    np.testing.assert_array_equal(got.view(np.uint64), np.ascontiguousarray(ref).view(np.uint64))


def test_spmm_k32_row_major_large(ctx):
    # The benchmark's SpMM shape (row-major 32-column block) on a Zipf matrix.
    m = G.make(G.ZIPF, 1 << 15, 1 << 15, seed=8, values=G.VALUES_NORMAL)
    k = 32
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, (m.cols, k))
    c0 = np.zeros((m.rows, k))
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, 4)
    got, (bb, cc, ldb, ldc) = run_spmm(ctx, H.OP_N, 1.0, a, b, c0, H.ROW_MAJOR)
    ref = O.std_spmm(0, 0, m.rows, m.cols, m.offsets, m.indices, m.values, 1.0, bb.copy(), 1, ldb, k,
                     cc.copy(), 1, ldc).reshape(m.rows, k)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("panel_bytes", [-1, 8 * 32 * 7, 8 * 32 * 1000, 1])
@pytest.mark.parametrize("op", [H.OP_N, H.OP_T])
def test_spmm_column_passes_bit_exact(ctx, panel_bytes, op):
    """SpMM in column passes (slices of B) keeps every C(i,j)'s terms in
    stored order: bit-exact with the oracle for any slice width, including
    one column per pass and the transposed (workspace) form."""
    m = G.make(G.ZIPF, 4000, 3000, seed=43, values=G.VALUES_NORMAL, zipf_max=900)
    k = 32
    nin, nout = (m.cols, m.rows) if op == H.OP_N else (m.rows, m.cols)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (nin, k))
    c0 = rng.uniform(-1, 1, (nout, k))
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, 4)
    print(f"spmm passes panel_bytes {panel_bytes} op {op}", flush=True)
    H.set_spmm_panel_bytes(ctx, panel_bytes)
    try:
        got, (bb, cc, ldb, ldc) = run_spmm(ctx, op, -0.75, a, b, c0, H.ROW_MAJOR)
    finally:
        H.set_spmm_panel_bytes(ctx, 0)
    ref = O.std_spmm(0, op, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), m.values, -0.75,
                     bb.copy(), 1, ldb, k, cc.copy(), 1, ldc).reshape(nout, k)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_spmm_column_passes_without_workspace_run_one_pass(ctx):
    """Column passes need a split table in the workspace; a call given none
    runs in one pass, with the same (bit-exact) result."""
    m = G.make(G.ZIPF, 2000, 1500, seed=44, values=G.VALUES_NORMAL, zipf_max=600)
    k = 32
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, (m.cols, k))
    c0 = rng.uniform(-1, 1, (m.rows, k))
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, 4)
    bb, cb = ctx.upload(b), ctx.upload(c0)
    bd = H.dense(bb, m.cols, k, k, H.ROW_MAJOR, H.F64, writable=False)
    cd = H.dense(cb, m.rows, k, k, H.ROW_MAJOR, H.F64)
    H.set_spmm_panel_bytes(ctx, 8 * k * 100)
    try:
        need = C.c_size_t()
        H.check(ctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
            ctx.handle, H.OP_N, C.byref(a.desc), C.byref(bd), C.byref(cd), C.byref(need)))
        assert need.value > 0  # the passes' split table
        H.check(ctx.lib.asc_sparse_hip_standard_spmm(ctx.handle, H.OP_N, H.Scalar(1.25, 0.0),
                                                     C.byref(a.desc), C.byref(bd), C.byref(cd),
                                                     None, 0))
    finally:
        H.set_spmm_panel_bytes(ctx, 0)
    got = ctx.download(cb, np.float64, m.rows * k).reshape(m.rows, k)
    ref = O.std_spmm(0, 0, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), m.values, 1.25,
                     b.copy(), 1, k, k, c0.copy(), 1, k).reshape(m.rows, k)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
