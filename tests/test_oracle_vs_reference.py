"""The oracle against the reference compiled from its own sources.

oracle/_ref/libref_driver.so is built by oracle/build_ref.sh from the
reference's serial Sparse BLAS/conversion sources.  On random inputs every
oracle function must match it bit for bit (indices, values, NaN-ness).
Skipped where the reference build is absent.
"""
import numpy as np
import pytest

import oracle_bind as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def random_csr(rng, rows, cols, density):
    mask = rng.random((rows, cols)) < density
    off = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int64)
    idx = np.nonzero(mask)[1].astype(np.int64)
    val = rng.standard_normal(idx.shape[0]) * np.exp(rng.standard_normal(idx.shape[0]) * 3)
    return off, idx, val


@pytest.mark.parametrize("seed", range(6))
def test_compat_spmv_matches_reference(seed):
    rng = np.random.default_rng(seed)
    rows, cols = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    off, idx, val = random_csr(rng, rows, cols, float(rng.uniform(0.01, 0.3)))
    x = rng.standard_normal(cols)
    y0 = rng.standard_normal(rows)
    for alpha, beta in ((1.0, 0.0), (-0.7, 1.3), (2.5, -0.0)):
        for dt in (np.float64, np.float32):
            a, b = O.compat_spmv(off, idx, val.astype(dt), alpha, x.astype(dt), beta, y0.astype(dt)), None
            b = O.ref_compat_spmv(off, idx, val.astype(dt), alpha, x.astype(dt), beta, y0.astype(dt))
            np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


def test_compat_spmv_matches_reference_on_zipf_rows():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "asc-cpp_amd"))
    from asc_sparse_hip import gen as G
    m = G.make(G.ZIPF, 20000, 20000, seed=9, values=G.VALUES_NORMAL)
    x = G.vector(m.cols, 4)
    a = O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, np.zeros(m.rows))
    b = O.ref_compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, np.zeros(m.rows))
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_standard_spmv_spmm_match_reference(seed, fmt, op):
    rng = np.random.default_rng(50 + seed)
    rows, cols = int(rng.integers(1, 120)), int(rng.integers(1, 120))
    off, idx, val = random_csr(rng, rows, cols, 0.15)
    if fmt == 1:
        off, idx, val = O.transpose(0, rows, cols, off, idx, val)
    nin, nout = (cols, rows) if op == 0 else (rows, cols)
    x = rng.standard_normal(nin)
    y0 = rng.standard_normal(nout)
    a = O.std_spmv(fmt, op, rows, cols, off, idx, val, -1.25, x, y0.copy())
    b = O.ref_std_spmv(fmt, op, rows, cols, off, idx, val, -1.25, x, y0.copy())
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))
    k = int(rng.integers(1, 9))
    for layout in (0, 1):
        bmat = rng.standard_normal((nin, k))
        cmat = rng.standard_normal((nout, k))
        if layout == 1:
            bb, cc, ldb, ldc = bmat.reshape(-1).copy(), cmat.reshape(-1).copy(), k, k
        else:
            bb, cc, ldb, ldc = bmat.T.reshape(-1).copy(), cmat.T.reshape(-1).copy(), nin, nout
        a = O.std_spmm(fmt, op, rows, cols, off, idx, val, 0.5, bb, layout, ldb, k, cc.copy(), layout, ldc)
        b = O.ref_std_spmm(fmt, op, rows, cols, off, idx, val, 0.5, bb, layout, ldb, k, cc.copy(), layout, ldc)
        np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("seed", range(4))
def test_conversions_match_reference(seed):
    rng = np.random.default_rng(80 + seed)
    rows, cols = int(rng.integers(0, 90)), int(rng.integers(0, 90))
    off, idx, val = random_csr(rng, rows, cols, 0.2) if rows and cols else (
        np.zeros(rows + 1, np.int64), np.zeros(0, np.int64), np.zeros(0))
    val[::7] = 0.0  # explicit zeros travel as stored entries
    coords, cval = O.compressed_to_coo(0, rows, off, idx, val)
    for fmt in (0, 1):
        a = O.coo_to_compressed(fmt, rows, cols, coords, cval)
        b = O.ref_coo_to_compressed(fmt, rows, cols, coords, cval)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    a = O.transpose(0, rows, cols, off, idx, val)
    b = O.ref_transpose(0, rows, cols, off, idx, val)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    a2 = O.transpose(1, rows, cols, *a)
    b2 = O.ref_transpose(1, rows, cols, *b)
    for u, v in zip(a2, b2):
        np.testing.assert_array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("dup,zero", [(1, 0), (1, 1), (0, 0), (0, 1)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_finalize_matches_reference(seed, dup, zero, dtype):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(0, 200))
    rows, cols = 9, 7
    r, c = rng.integers(0, rows, n), rng.integers(0, cols, n)
    if dup == 0:  # mostly unique coordinates so kReject sometimes succeeds
        flat = rng.choice(rows * cols, size=min(n, rows * cols), replace=seed % 2 == 0)
        r, c = flat // cols, flat % cols
        n = flat.shape[0]
    # small integers (exact zeros after cancellation) mixed with normals,
    # whose sums depend on the order of addition
    vals = np.where(rng.random(n) < 0.5, rng.integers(-2, 3, n), rng.standard_normal(n)).astype(dtype)
    vals[::11] = np.nan
    coords = np.stack([r, c], 1).reshape(-1)
    a = O.finalize(coords, vals, dup, zero)
    b = O.ref_finalize(rows, cols, coords, vals, dup, zero)
    assert a[0] == b[0]
    if a[0] == 0:
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(np.isnan(a[2]), np.isnan(b[2]))
        np.testing.assert_array_equal(np.nan_to_num(a[2]), np.nan_to_num(b[2]))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("op", ["dot", "axpy", "gather", "gather_zero", "scatter"])
@pytest.mark.parametrize("seed", range(3))
def test_level1_matches_reference(op, dtype, seed):
    rng = np.random.default_rng(300 + seed)
    extent = int(rng.integers(1, 400))
    inc = int(rng.integers(1, 4))
    nnz = int(rng.integers(0, extent + 1))
    idx = np.sort(rng.choice(extent, nnz, replace=False)).astype(np.int64)
    v = rng.standard_normal(nnz).astype(dtype)
    x = rng.standard_normal((extent - 1) * inc + 1).astype(dtype)
    alpha = [0.0, 1.5, -0.25][seed]
    a = O.level1(op, idx, v, x, inc, alpha)
    rc, dot, rv, rx = O.ref_level1(op, idx, v, x, extent, inc, alpha)
    assert rc == 0
    if op == "dot":
        assert np.asarray([a[0]], dtype).tobytes() == np.asarray([dot], dtype).tobytes()
    np.testing.assert_array_equal(a[1].view(np.uint8), rv.view(np.uint8))
    np.testing.assert_array_equal(a[2].view(np.uint8), rx.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("op", [0, 1, 2])
@pytest.mark.parametrize("lower,unit", [(True, False), (False, False), (True, True), (False, True)])
def test_triangular_matches_reference(dtype, fmt, op, lower, unit):
    rng = np.random.default_rng(17 + fmt * 7 + op * 3 + lower * 2 + unit)
    n, k = 40, 3
    dense = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.15)
    dense = np.tril(dense, -1) if lower else np.triu(dense, 1)
    dense[np.arange(n), np.arange(n)] = rng.uniform(1.0, 2.0, n) * rng.choice([-1, 1], n)
    dense = dense.astype(dtype)
    src = dense if fmt == 0 else dense.T  # CSC of A = CSR of A^T
    off = np.concatenate([[0], np.cumsum((src != 0).sum(1))])
    idx = np.concatenate([np.nonzero(r)[0] for r in src])
    val = src[src != 0].astype(dtype)
    b = rng.standard_normal((n, k)).astype(dtype)
    for alpha in (1.0, -0.5, 0.0):
        rc, ref = O.ref_trsv(fmt, op, lower, unit, n, off, idx, val, alpha, b)
        assert rc == 0
        got = O.trsv(fmt, op, lower, unit, n, off, idx, val, alpha, b)
        np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))
