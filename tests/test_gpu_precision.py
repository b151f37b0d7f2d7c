"""Reduced-precision storage: FP32 / BF16 / FP16 values, FP32 x and y, FP32
accumulation (plain or compensated TwoSum), error against the FP64 oracle on
the original FP64 values:  err = ||y - y64||_inf / || |A||x| ||_inf.

Workload: the precision case (32 nnz/row: 16 within +-64 of the diagonal and
16 uniform; values random sign * exp(N(0,2))), at 2^16 rows here.
"""
import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu

# Storage rounding bounds the error: |A|'s relative rounding (2^-24, 2^-9,
# 2^-12) plus x's FP32 rounding plus FP32 accumulation over 32 terms.
TOL = {H.F32: 4e-6, H.BF16: 1e-2, H.F16: 1.5e-3}


@pytest.fixture(scope="module")
def problem():
    m = G.precision_matrix(1 << 16)
    x64 = G.vector(m.cols, 33)
    y64 = O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x64, 0.0, np.zeros(m.rows))
    absax = O.abs_spmv(m.offsets, m.indices, m.values, x64)
    return m, x64, y64, absax


def f16_scale(values):
    # Power of two keeping max|v| * scale below FP16's largest finite value.
    vmax = float(np.max(np.abs(values)))
    scale = 1.0
    while vmax * scale > 32768.0:
        scale *= 0.5
    return scale


@pytest.mark.parametrize("kind", [H.F32, H.BF16, H.F16])
@pytest.mark.parametrize("acc", [H.ACC_PLAIN, H.ACC_TWOSUM])
@pytest.mark.parametrize("alg", [H.ALG_MERGE, H.ALG_VECTOR[8], H.ALG_VECTOR[32]])
def test_error_against_fp64_oracle(ctx, problem, kind, acc, alg):
    m, x64, y64, absax = problem
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    scale = f16_scale(m.values) if kind == H.F16 else 1.0
    ak = H.convert_values(ctx, a64, kind, scale)
    xb = ctx.upload(x64.astype(np.float32))
    yb = ctx.alloc(4 * m.rows)
    H.csr_spmv(ctx, 1.0 / scale, ak, H.vector(xb, m.cols, H.F32, writable=False), 0.0,
               H.vector(yb, m.rows, H.F32), alg=alg, acc=acc)
    y = ctx.download(yb, np.float32, m.rows).astype(np.float64)
    err = np.max(np.abs(y - y64)) / np.max(absax)
    print(f"kind {kind} acc {acc} alg {H.ALG_NAMES[alg]} scale {scale} err {err:.3e}", flush=True)
    assert np.isfinite(y).all()
    assert err <= TOL[kind]


def test_twosum_not_worse_than_plain_for_fp32(ctx, problem):
    m, x64, y64, absax = problem
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    a32 = H.convert_values(ctx, a64, H.F32)
    xb = ctx.upload(x64.astype(np.float32))
    errs = {}
    for acc in (H.ACC_PLAIN, H.ACC_TWOSUM):
        yb = ctx.alloc(4 * m.rows)
        H.csr_spmv(ctx, 1.0, a32, H.vector(xb, m.cols, H.F32, writable=False), 0.0,
                   H.vector(yb, m.rows, H.F32), alg=H.ALG_MERGE, acc=acc)
        y = ctx.download(yb, np.float32, m.rows).astype(np.float64)
        # Against the exact product of the stored FP32 operands.
        v32 = m.values.astype(np.float32).astype(np.float64)
        x32 = x64.astype(np.float32).astype(np.float64)
        yex = O.compat_spmv(m.offsets, m.indices, v32, 1.0, x32, 0.0, np.zeros(m.rows))
        errs[acc] = np.max(np.abs(y - yex) / np.maximum(O.abs_spmv(m.offsets, m.indices, v32, x32), 1e-300))
    assert errs[H.ACC_TWOSUM] <= errs[H.ACC_PLAIN]
    assert errs[H.ACC_TWOSUM] <= 2.0 ** -23  # one final FP32 rounding


def test_value_conversion_round_to_nearest_even(ctx):
    # Ties and near-ties against numpy's correctly rounded casts (FP32, FP16)
    # and an exact integer construction for BF16.
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(5000) * np.exp(rng.standard_normal(5000) * 4),
                        [0.0, -0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 65504.0, 65520.0, 1e-8,
                         3.4e38, 1e300, -1e300]])
    n = v.shape[0]
    off = np.array([0, n], np.int64)
    idx = np.arange(n, dtype=np.int64)
    a = H.clone_compressed(ctx, H.CSR, 1, n, off, idx, v, H.F64)
    got32 = H.download_compressed(ctx, H.convert_values(ctx, a, H.F32))[2]
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(got32.view(np.uint32), v.astype(np.float32).view(np.uint32))
        got16 = H.download_compressed(ctx, H.convert_values(ctx, a, H.F16))[2]
        np.testing.assert_array_equal(got16.view(np.uint16), v.astype(np.float16).view(np.uint16))
    gotb = H.download_compressed(ctx, H.convert_values(ctx, a, H.BF16))[2]
    np.testing.assert_array_equal(gotb, bf16_rne(v))


def bf16_rne(v):
    """Correctly rounded FP64 -> BF16 bit patterns via exact rationals."""
    from fractions import Fraction
    import math
    out = np.empty(v.shape[0], np.uint16)
    for i, d in enumerate(v.tolist()):
        sign = 0x8000 if math.copysign(1.0, d) < 0 else 0
        a = abs(d)
        if a == 0.0:
            out[i] = sign
            continue
        e = math.frexp(a)[1] - 1  # a in [2^e, 2^(e+1))
        e = max(e, -126)
        q = Fraction(a) / Fraction(2) ** (e - 7)  # 7 fraction bits
        n = math.floor(q)
        r = q - n
        if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
            n += 1
        val = Fraction(n) * Fraction(2) ** (e - 7)
        if val >= Fraction(2) ** 128:
            out[i] = sign | 0x7F80
            continue
        f = float(val)
        bits = np.array([f], np.float32).view(np.uint32)[0]
        out[i] = sign | (int(bits) >> 16)
    return out


@pytest.mark.parametrize("alg", [H.ALG_VECTOR[8], H.ALG_VECTOR[16]])
def test_paired_loads_and_misaligned_values(ctx, alg):
    # 16-bit values start on a 4-byte boundary in a clone, so the vector
    # kernel loads them in pairs; the same values rebound one element off
    # that boundary take the per-entry loop.  Both sum the same products
    # (only the order differs), within the FP32 plain-mode bound.
    m = G.make(G.ZIPF, 1 << 14, 1 << 14, seed=31, values=G.VALUES_NORMAL)
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    abf = H.convert_values(ctx, a64, H.BF16)
    bits = ctx.download(H._ptr_buf(ctx, abf.desc.values), np.uint16, m.nnz)
    src = ctx.upload(bits)
    shifted = ctx.alloc(2 * (m.nnz + 2))
    ctx.copy(shifted.ptr + 2, src.ptr, 2 * m.nnz)
    ctx.synchronize()
    odd = H.rebind_values(abf, shifted.ptr + 2)
    x = G.vector(m.cols, 32).astype(np.float32)
    xb = ctx.upload(x)
    xd = H.vector(xb, m.cols, H.F32, writable=False)
    ys = []
    for a in (abf, odd):
        yb = ctx.alloc(4 * m.rows)
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, H.vector(yb, m.rows, H.F32), alg=alg)
        ys.append(ctx.download(yb, np.float32, m.rows).astype(np.float64))
    vals = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    absax = O.abs_spmv(m.offsets, m.indices, vals, x.astype(np.float64))
    lens = np.maximum(np.diff(m.offsets), 1)
    bound = 2 * 4.0 * lens * 2.0 ** -24 * absax
    bad = np.nonzero(~(np.abs(ys[0] - ys[1]) <= bound))[0]
    assert bad.size == 0, f"{bad.size} rows differ beyond the bound, first {bad[:5]}"
    ref = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, vals, 1.0, x.astype(np.float64), 0.0, ref)
    for y in ys:
        assert np.all(np.abs(y - ref) <= 4.0 * lens * 2.0 ** -24 * absax + 1e-300)
