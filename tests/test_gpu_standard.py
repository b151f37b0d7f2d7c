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


def run_spmm(ctx, op, alpha, a, b, c0, layout, alg=H.ALG_ORDERED, ws=None):
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
                    H.dense(cbuf, nc, k, ldc, layout, kind), ws=ws, alg=alg)
    out = ctx.download(cbuf, dt, nc * k)
    return (out.reshape(nc, k) if layout == H.ROW_MAJOR else out.reshape(k, nc).T), (bb, cc, ldb, ldc)


def op_rows(op, rows, cols, off, idx, val):
    """op(A) of a CSR matrix as (outer, offsets, indices, values) rows."""
    if op == H.OP_N:
        return rows, off, idx, val
    t_off, t_idx, t_val = O.transpose(0, rows, cols, off, idx, val)
    return cols, t_off, t_idx, t_val


def plain_spmm_bound(op, rows, cols, off, idx, val, alpha, b, c0, eps):
    """Plain-form bound on |C - C_serial|: 4*(len_i+1)*eps*(|alpha|(|A||B|)_ij + |C0_ij|),
    with |A||B| summed by a sparse product of absolute values (a magnitude, so
    its own rounding is immaterial)."""
    import scipy.sparse as sp
    n, o, i, v = op_rows(op, rows, cols, off, np.asarray(idx, np.int64), val)
    abs_a = sp.csr_matrix((np.abs(v.astype(np.float64)), i, o), shape=(n, b.shape[0]))
    lens = np.diff(np.asarray(o, np.int64))
    mag = abs(alpha) * (abs_a @ np.abs(b.astype(np.float64))) + np.abs(c0.astype(np.float64))
    return 4 * (lens[:, None] + 1) * eps * mag


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
@pytest.mark.parametrize("k", [1, 4, 5, 32, 70])
@pytest.mark.parametrize("alg", [H.ALG_ORDERED, H.ALG_AUTO])
def test_standard_spmm(ctx, fmt, op, layout, k, alg):
    """Ordered form: bit-exact with the oracle.  Plain form: within the plain
    bound where its kernel applies (row-major, k in {4, 32}); every other
    shape runs the ordered kernels, so AUTO is bit-exact there too."""
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
    print(f"spmm fmt {fmt} op {op} layout {layout} k {k} alg {alg}", flush=True)
    got, (bb, cc, ldb, ldc) = run_spmm(ctx, op, 1.5, a, b, c0, layout, alg)
    ref = O.std_spmm(0 if fmt == H.CSR else 1, op, rows, cols, off, idx, val, 1.5, bb.copy(), layout,
                     ldb, k, cc.copy(), layout, ldc)
    ref = ref.reshape(nout, k) if layout == H.ROW_MAJOR else ref.reshape(k, nout).T
    plain = alg == H.ALG_AUTO and layout == H.ROW_MAJOR and k % 4 == 0 and k <= 32
    if not plain:
        np.testing.assert_array_equal(got.view(np.uint64), np.ascontiguousarray(ref).view(np.uint64))
        return
    # op(A) in CSR row form, for the bound
    if fmt == H.CSR:
        bound = plain_spmm_bound(op, rows, cols, off, idx, val, 1.5, b, c0, EPS[np.float64])
    else:
        t_off, t_idx, t_val = O.transpose(0, cols, rows, off, idx, val)  # CSC -> CSR of A
        bound = plain_spmm_bound(op, rows, cols, t_off, t_idx, t_val, 1.5, b, c0, EPS[np.float64])
    err = np.abs(got - ref)
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))
    assert np.any(err > 0) or k == 4  # the plain kernel ran (FMA rounding differs somewhere)


@pytest.mark.parametrize("alg", [H.ALG_ORDERED, H.ALG_AUTO])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmm_k32_row_major_large(ctx, alg, dtype):
    # The benchmark's SpMM shape (row-major 32-column block) on a Zipf matrix
    # with rows from 1 to 10^4 entries (both plain row classes).
    m = G.make(G.ZIPF, 1 << 15, 1 << 15, seed=8, values=G.VALUES_NORMAL)
    k = 32
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, (m.cols, k)).astype(dtype)
    c0 = rng.uniform(-1, 1, (m.rows, k)).astype(dtype)
    val = m.values.astype(dtype)
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, val, 4)
    got, (bb, cc, ldb, ldc) = run_spmm(ctx, H.OP_N, -0.5, a, b, c0, H.ROW_MAJOR, alg)
    ref = O.std_spmm(0, 0, m.rows, m.cols, m.offsets, m.indices, val, -0.5, bb.copy(), 1, ldb, k,
                     cc.copy(), 1, ldc).reshape(m.rows, k)
    if alg == H.ALG_ORDERED:
        np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))
    else:
        bound = plain_spmm_bound(H.OP_N, m.rows, m.cols, m.offsets, m.indices, val, -0.5, b, c0,
                                 EPS[dtype])
        err = np.abs(got.astype(np.float64) - ref.astype(np.float64))
        assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))


@pytest.mark.parametrize("index_bytes", [4, 8])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("k,ld_pad", [(4, 0), (8, 4), (12, 0), (16, 8), (28, 4), (32, 0)])
def test_spmm_plain_shapes(ctx, index_bytes, dtype, k, ld_pad):
    """The plain kernel over int32 and int64 indices, FP64 and FP32, k a
    multiple of 4 up to 32, and leading dimensions beyond k (row-major B and
    C with ld = k + ld_pad): within the plain bound, the padding columns of C
    untouched, the result identical on a second run."""
    m = G.make(G.ZIPF, 2500, 1800, seed=47, values=G.VALUES_NORMAL, zipf_max=700)
    val = m.values.astype(dtype)
    rng = np.random.default_rng(k + ld_pad)
    ld = k + ld_pad
    b = rng.uniform(-1, 1, (m.cols, ld)).astype(dtype)
    c0 = rng.uniform(-1, 1, (m.rows, ld)).astype(dtype)
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, val, index_bytes)
    kind = KIND[dtype]
    bb = ctx.upload(b.reshape(-1))
    outs = []
    print(f"spmm plain shapes {index_bytes} {dtype.__name__} k {k} ld {ld}", flush=True)
    for _ in range(2):
        cb = ctx.upload(c0.reshape(-1))
        H.standard_spmm(ctx, H.OP_N, -1.25, a, H.dense(bb, m.cols, k, ld, H.ROW_MAJOR, kind, writable=False),
                        H.dense(cb, m.rows, k, ld, H.ROW_MAJOR, kind), alg=H.ALG_AUTO)
        outs.append(ctx.download(cb, dtype, m.rows * ld).reshape(m.rows, ld))
    got = outs[0]
    np.testing.assert_array_equal(outs[1].view(np.uint8), got.view(np.uint8))  # repeatable
    np.testing.assert_array_equal(got[:, k:], c0[:, k:])  # padding untouched
    ref = O.std_spmm(0, 0, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), val, -1.25,
                     b.reshape(-1).copy(), 1, ld, k, c0.reshape(-1).copy(), 1, ld).reshape(m.rows, ld)
    bound = plain_spmm_bound(H.OP_N, m.rows, m.cols, m.offsets, m.indices, val, -1.25, b[:, :k],
                             c0[:, :k], EPS[dtype])
    err = np.abs(got[:, :k].astype(np.float64) - ref[:, :k].astype(np.float64))
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))


def test_spmm_plain_falls_back_to_ordered_on_misaligned_blocks(ctx):
    """B (or C) not aligned to 4 elements, or a leading dimension that is not
    a multiple of 4, runs the ordered kernels: bit-exact with the oracle."""
    m = G.make(G.ZIPF, 1500, 1200, seed=48, values=G.VALUES_NORMAL, zipf_max=500)
    a = device_matrix(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, 4)
    rng = np.random.default_rng(9)
    k = 8
    for case, (b_off, ld) in {"B one element off": (1, 8), "ld 9": (0, 9)}.items():
        b = rng.uniform(-1, 1, (m.cols, ld))
        c0 = rng.uniform(-1, 1, (m.rows, ld))
        bflat = np.concatenate([np.zeros(b_off), b.reshape(-1)])
        bb, cb = ctx.upload(bflat), ctx.upload(c0.reshape(-1))
        print(f"spmm fallback: {case}", flush=True)
        H.standard_spmm(ctx, H.OP_N, 0.5, a,
                        H.dense(bb.ptr + 8 * b_off, m.cols, k, ld, H.ROW_MAJOR, H.F64, writable=False),
                        H.dense(cb, m.rows, k, ld, H.ROW_MAJOR, H.F64), alg=H.ALG_AUTO)
        got = ctx.download(cb, np.float64, m.rows * ld).reshape(m.rows, ld)
        ref = O.std_spmm(0, 0, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), m.values, 0.5,
                         b.reshape(-1).copy(), 1, ld, k, c0.reshape(-1).copy(), 1, ld).reshape(m.rows, ld)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64), err_msg=case)


# 256 MiB slices of B: FP64, 32 columns -> 2^20 rows of B per pass.
PASS_ROWS = (256 << 20) // (8 * 32)


@pytest.mark.parametrize("op", [H.OP_N, H.OP_T])
def test_spmm_plain_column_passes(ctx, op):
    """B larger than one 256 MiB slice (2.5 * 2^20 rows, three passes) with the
    passes' split table in the workspace, and the same call without a
    workspace (one pass): both within the plain bound of the serial result,
    for A and for its transpose; the ordered form bit for bit with both
    workspaces."""
    k, short, long_ = 32, 1 << 14, int(2.5 * PASS_ROWS)
    rng = np.random.default_rng(11)
    if op == H.OP_N:   # 2^14 x 2.6M, ~48 entries a row spread over every pass
        rows, cols = short, long_
        lens = rng.integers(0, 97, rows)
        lens[:64] = 3000  # long rows too
    else:              # 2.6M x 2^14, transposed into 2^14 rows of op(A)
        rows, cols = long_, short
        lens = rng.integers(0, 2, rows)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.int32)
    val = rng.standard_normal(len(idx))
    nin, nout = (cols, rows) if op == H.OP_N else (rows, cols)
    b = rng.uniform(-1, 1, (nin, k))
    c0 = rng.uniform(-1, 1, (nout, k))
    a = device_matrix(ctx, H.CSR, rows, cols, off, idx, val, 4)
    ref = O.std_spmm(0, op, rows, cols, off, idx.astype(np.int64), val, 0.75, b.reshape(-1).copy(), 1,
                     k, k, c0.reshape(-1).copy(), 1, k).reshape(nout, k)
    bound = plain_spmm_bound(op, rows, cols, off, idx, val, 0.75, b, c0, EPS[np.float64])
    bb, cb = ctx.upload(b.reshape(-1)), ctx.upload(c0.reshape(-1))
    bd = H.dense(bb, nin, k, k, H.ROW_MAJOR, H.F64, writable=False)
    cd = H.dense(cb, nout, k, k, H.ROW_MAJOR, H.F64)
    need = C.c_size_t()
    H.check(ctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
        ctx.handle, op, C.byref(a.desc), C.byref(bd), C.byref(cd), None, C.byref(need)))
    # The transposed operand's part (op T) without the passes' split table:
    # the query for a column-major B, which no column pass serves.
    prefix = C.c_size_t()
    bd_cm = H.dense(bb, nin, k, nin, H.COLUMN_MAJOR, H.F64, writable=False)
    H.check(ctx.lib.asc_sparse_hip_standard_spmm_workspace_size(
        ctx.handle, op, C.byref(a.desc), C.byref(bd_cm), C.byref(cd), None, C.byref(prefix)))
    assert need.value >= prefix.value + 2 * nout * 4  # two split rows of uint32
    for label, give in (("passes", need.value), ("one pass", prefix.value)):
        print(f"spmm plain passes op {op} {label}", flush=True)
        H.check(ctx.lib.asc_sparse_hip_copy(ctx.handle, cb.ptr, c0.ctypes.data, c0.nbytes))
        ctx.synchronize()
        ws = H.Workspace(ctx)
        p, n = ws.get(give)
        H.check(ctx.lib.asc_sparse_hip_standard_spmm(ctx.handle, op, H.Scalar(0.75, 0.0),
                                                     C.byref(a.desc), C.byref(bd), C.byref(cd),
                                                     p, n, None))
        got = ctx.download(cb, np.float64, nout * k).reshape(nout, k)
        err = np.abs(got - ref)
        assert np.all(err <= bound), (label, float(np.max(err / np.maximum(bound, 1e-300))))
        # The ordered form takes the same passes (each C(i,j)'s chain goes on
        # pass after pass in stored order): the serial result bit for bit.
        H.check(ctx.lib.asc_sparse_hip_copy(ctx.handle, cb.ptr, c0.ctypes.data, c0.nbytes))
        ctx.synchronize()
        H.check(ctx.lib.asc_sparse_hip_standard_spmm(ctx.handle, op, H.Scalar(0.75, 0.0),
                                                     C.byref(a.desc), C.byref(bd), C.byref(cd),
                                                     p, n, C.byref(H.options(H.ALG_ORDERED))))
        got = ctx.download(cb, np.float64, nout * k).reshape(nout, k)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
