"""GPU parity of the level-scheduled sparse triangular solve (blas.h:1760-1880)
through the C ABI: the golden vectors of standard_blas_test.cc:335-430
(CSR lower, CSC conjugate transpose, several right-hand sides, alpha == 0
over NaN values); random triangular matrices in CSR and CSC, every op,
lower/upper, unit/non-unit, FP64/FP32, one and three right-hand sides in
both layouts, bit-exact against the oracle (itself pinned to the
reference's solver); the reference's view checks (outside-triangle entry,
singular non-unit diagonal, non-square); a long dependency chain.
"""
import json
import os

import numpy as np
import pytest

import asc_sparse_hip as H
import oracle_bind as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
KIND = {np.float64: H.F64, np.float32: H.F32}
FMT = {0: H.CSR, 1: H.CSC}
OPS = {0: H.OP_N, 1: H.OP_T, 2: H.OP_H}


def solve(ctx, fmt, op, lower, unit, n, off, idx, val, alpha, b, layout=H.ROW_MAJOR):
    dtype = val.dtype.type
    kind = KIND[dtype]
    a = H.clone_compressed(ctx, FMT[fmt], n, n, np.asarray(off), np.asarray(idx, np.int64), val, kind)
    plan = H.triangular_analyze(ctx, OPS[op], H.LOWER if lower else H.UPPER,
                                H.UNIT if unit else H.NON_UNIT, a)
    b2 = np.asarray(b, dtype).reshape(n, -1)
    k = b2.shape[1]
    host = b2 if layout == H.ROW_MAJOR else np.ascontiguousarray(b2.T)
    buf = ctx.upload(np.ascontiguousarray(host))
    ld = k if layout == H.ROW_MAJOR else n
    rhs = H.dense(buf, n, k, max(ld, 1), layout, kind)
    H.triangular_solve(ctx, OPS[op], alpha, a, rhs, plan)
    out = ctx.download(buf, dtype, n * k)
    out = out.reshape(n, k) if layout == H.ROW_MAJOR else out.reshape(k, n).T
    return np.ascontiguousarray(out.reshape(np.shape(b))), plan


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_golden_vectors(ctx, dtype):
    g = GOLD["triangular"]
    v = np.array(g["values"], dtype)
    x, _ = solve(ctx, 0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], v, 1.0, g["lower_n_rhs"])
    np.testing.assert_array_equal(x, g["lower_n_solution"])
    x, _ = solve(ctx, 1, 2, True, False, 2, g["csc_offsets"], g["csc_indices"], v, 1.0,
                 g["csc_lower_h_rhs"])
    np.testing.assert_array_equal(x, g["csc_lower_h_solution"])
    b = np.array(g["multiple_rhs_row_major"]).reshape(2, 2)
    for layout in (H.ROW_MAJOR, H.COLUMN_MAJOR):
        x, _ = solve(ctx, 0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], v, 1.0, b, layout)
        np.testing.assert_array_equal(x.reshape(-1), g["multiple_solution"])
    zv = np.array([float(t) for t in g["alpha_zero_values"]], dtype)
    # NaN off-diagonal value: the view checks pass (diagonal 2, 3 nonzero)
    x, _ = solve(ctx, 0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], zv, 0.0, g["alpha_zero_rhs"])
    np.testing.assert_array_equal(x, g["alpha_zero_solution"])


def triangular(rng, n, lower, dtype, density=0.1):
    dense = rng.standard_normal((n, n)) * (rng.random((n, n)) < density)
    dense = np.tril(dense, -1) if lower else np.triu(dense, 1)
    dense[np.arange(n), np.arange(n)] = rng.uniform(1.0, 2.0, n) * rng.choice([-1, 1], n)
    return dense.astype(dtype)


def compressed(dense, fmt):
    src = dense if fmt == 0 else dense.T
    off = np.concatenate([[0], np.cumsum((src != 0).sum(1))])
    idx = np.concatenate([np.nonzero(r)[0] for r in src]).astype(np.int64)
    return off, idx, src[src != 0]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("op", [0, 1, 2])
@pytest.mark.parametrize("lower,unit", [(True, False), (False, False), (True, True), (False, True)])
def test_random_against_oracle(ctx, dtype, fmt, op, lower, unit):
    rng = np.random.default_rng(100 + 17 * fmt + 5 * op + 3 * lower + unit)
    n = 300
    dense = triangular(rng, n, lower, dtype)
    off, idx, val = compressed(dense, fmt)
    print(f"trsv {dtype.__name__} fmt {fmt} op {op} lower {lower} unit {unit}", flush=True)
    for k, layout in ((1, H.ROW_MAJOR), (3, H.ROW_MAJOR), (3, H.COLUMN_MAJOR)):
        b = rng.standard_normal((n, k) if k > 1 else n).astype(dtype)
        for alpha in (1.0, -0.75):
            ref = O.trsv(fmt, op, lower, unit, n, off, idx, val, alpha, b)
            got, plan = solve(ctx, fmt, op, lower, unit, n, off, idx, val, alpha, b, layout)
            np.testing.assert_array_equal(got.view(np.uint8), np.ascontiguousarray(ref).view(np.uint8))
            assert 1 <= plan.info.levels <= n


def test_view_checks(ctx):
    # entry outside the declared triangle
    off, idx, val = np.array([0, 2, 3]), np.array([0, 1, 1]), np.array([1.0, 2.0, 3.0])
    a = H.clone_compressed(ctx, H.CSR, 2, 2, off, idx, val, H.F64)
    with pytest.raises(H.AscHipError) as e:
        H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, a)
    assert e.value.code == H.INVALID_ARGUMENT
    H.triangular_analyze(ctx, H.OP_N, H.UPPER, H.NON_UNIT, a)  # the same entries are upper
    # singular non-unit (blas_test.cc-style: diagonal {1, 0})
    s = H.clone_compressed(ctx, H.CSR, 2, 2, np.array([0, 1, 2]), np.array([0, 1]), np.array([1.0, 0.0]),
                           H.F64)
    with pytest.raises(H.AscHipError) as e:
        H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, s)
    assert e.value.code == H.INVALID_ARGUMENT
    H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.UNIT, s)  # unit diagonal ignores it
    r = H.clone_compressed(ctx, H.CSR, 2, 3, np.array([0, 1, 2]), np.array([0, 1]), np.ones(2), H.F64)
    with pytest.raises(H.AscHipError) as e:
        H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, r)
    assert e.value.code == H.SHAPE


def test_long_chain_and_wide_level(ctx):
    """A bidiagonal matrix (one row per level, n levels) and a diagonal one
    (one level of n rows)."""
    n = 2000
    rng = np.random.default_rng(11)
    dense = np.diag(rng.uniform(1, 2, n)) + np.diag(rng.uniform(-0.5, 0.5, n - 1), -1)
    off, idx, val = compressed(dense, 0)
    b = rng.standard_normal(n)
    ref = O.trsv(0, 0, True, False, n, off, idx, val, 1.0, b)
    got, plan = solve(ctx, 0, 0, True, False, n, off, idx, val, 1.0, b)
    assert plan.info.levels == n and plan.info.max_level_rows == 1
    np.testing.assert_array_equal(got.view(np.uint8), ref.view(np.uint8))
    diag = np.diag(rng.uniform(1, 2, n))
    off, idx, val = compressed(diag, 0)
    got, plan = solve(ctx, 0, 1, True, False, n, off, idx, val, 2.0, b)
    assert plan.info.levels == 1 and plan.info.max_level_rows == n
    np.testing.assert_array_equal(got, O.trsv(0, 1, True, False, n, off, idx, val, 2.0, b))
