"""GPU parity of the plain SpMM over a stream-staged column-panel analysis
(ASC_HIP_ALG_PANEL): the analysis's listed rows go segment by segment (one
segment per row and panel, B gathered from the panel's slice), the other rows
through the plain kernel.

Every result is checked against the serial oracle (the reference's
standardized SpMM, blas.h:1692-1750) within the plain-form bound
4*(len_i+1)*eps*(|alpha|(|A||B|)_ij + |C0_ij|), and against itself on a
second run (the combine adds the panels in a fixed order).  Rows the analysis
does not list take the plain kernel and equal its output bit for bit.
"""
import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O
from test_gpu_standard import EPS, KIND, plain_spmm_bound

pytestmark = pytest.mark.gpu


def spmm(ctx, a, bbuf, b_shape, c0, k, ld, alpha, alg, dtype):
    kind = KIND[dtype]
    cb = ctx.upload(c0.reshape(-1))
    H.standard_spmm(ctx, H.OP_N, alpha, a, H.dense(bbuf, b_shape, k, ld, H.ROW_MAJOR, kind, writable=False),
                    H.dense(cb, c0.shape[0], k, ld, H.ROW_MAJOR, kind), alg=alg)
    return ctx.download(cb, dtype, c0.size).reshape(c0.shape)


CASES = [
    # (rows, cols, zipf_max, long_row_min, panel_columns, k, ld, dtype, index_bytes)
    (1 << 15, 1 << 15, 10_000, 1024, 0, 32, 32, np.float64, 4),
    (1 << 15, 1 << 15, 10_000, 64, 2048, 32, 32, np.float64, 8),
    (1 << 15, 1 << 15, 10_000, 1, 0, 16, 20, np.float32, 4),
    (5000, 70_000, 3000, 16, 97, 20, 24, np.float64, 4),
    (5000, 70_000, 3000, 128, 4096, 4, 4, np.float32, 8),
    (20_000, 20_000, 12, 8, 1000, 28, 28, np.float64, 4),
]


@pytest.mark.parametrize("rows,cols,zmax,long_min,pcols,k,ld,dtype,ib", CASES)
def test_spmm_panel_parity(ctx, rows, cols, zmax, long_min, pcols, k, ld, dtype, ib):
    m = G.make(G.ZIPF, rows, cols, seed=rows % 1013 + long_min, values=G.VALUES_NORMAL, zipf_max=zmax)
    val = m.values.astype(dtype)
    rng = np.random.default_rng(k * 7 + long_min)
    b = rng.uniform(-1, 1, (cols, ld)).astype(dtype)
    c0 = rng.uniform(-1, 1, (rows, ld)).astype(dtype)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, m.offsets, m.indices, val, KIND[dtype], ib)
    bbuf = ctx.upload(b.reshape(-1))
    # The plain form before any analysis (AUTO may take the panel form after).
    plain = spmm(ctx, a, bbuf, cols, c0, k, ld, -0.75, H.ALG_AUTO, dtype)
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(panel_columns=pcols, long_row_min=long_min,
                                                      staging=H.STAGING_STREAM))
    print(f"spmm panel rows {rows} cols {cols} long_min {long_min} panel {pcols} k {k} ld {ld} "
          f"{dtype.__name__} ib {ib} nnz {m.nnz}", flush=True)
    got = spmm(ctx, a, bbuf, cols, c0, k, ld, -0.75, H.ALG_PANEL, dtype)
    again = spmm(ctx, a, bbuf, cols, c0, k, ld, -0.75, H.ALG_PANEL, dtype)
    np.testing.assert_array_equal(again.view(np.uint8), got.view(np.uint8))  # repeatable
    np.testing.assert_array_equal(got[:, k:], c0[:, k:])  # padding columns untouched
    ref = O.std_spmm(0, 0, rows, cols, m.offsets, m.indices.astype(np.int64), val, -0.75,
                     b.reshape(-1).copy(), 1, ld, k, c0.reshape(-1).copy(), 1, ld).reshape(rows, ld)
    bound = plain_spmm_bound(H.OP_N, rows, cols, m.offsets, m.indices, val, -0.75, b[:, :k], c0[:, :k],
                             EPS[dtype])
    err = np.abs(got[:, :k].astype(np.float64) - ref[:, :k].astype(np.float64))
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))
    # Rows below long_row_min ran the plain kernel: the same bits as the
    # plain form without the analysis.
    short = np.diff(m.offsets) < long_min
    np.testing.assert_array_equal(got[short].view(np.uint8), plain[short].view(np.uint8))
    del plan


def _spmm_raw(ctx, a, bbuf, rows, cols, c0, k, alg, ws, ws_bytes):
    """One SpMM through the C ABI with the caller's workspace pointer and size."""
    cb = ctx.upload(c0.reshape(-1))
    bd = H.dense(bbuf, cols, k, k, H.ROW_MAJOR, H.F64, writable=False)
    cd = H.dense(cb, rows, k, k, H.ROW_MAJOR, H.F64)
    H.check(ctx.lib.asc_sparse_hip_standard_spmm(ctx.handle, H.OP_N, H.Scalar(1.25, 0.0),
                                                 H.C.byref(a.desc), H.C.byref(bd), H.C.byref(cd),
                                                 ws, ws_bytes, H.C.byref(H.options(alg))))
    return ctx.download(cb, np.float64, c0.size).reshape(c0.shape)


def test_spmm_auto_takes_panel_form_when_rows_are_long(ctx):
    # AUTO runs the panel form on a stream-staged analysis whose listed rows
    # hold at least 8 entries per panel (here 1024 >= 8 * 8 panels): the
    # same bits as ALG_PANEL.  A workspace too small for its partial sums
    # runs the plain form instead, and so does an analysis listing every row.
    rows = cols = 1 << 15
    m = G.make(G.ZIPF, rows, cols, seed=17, values=G.VALUES_NORMAL, zipf_max=10_000)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, m.offsets, m.indices, m.values, H.F64, 4)
    rng = np.random.default_rng(23)
    k = 32
    b = rng.uniform(-1, 1, (cols, k))
    c0 = rng.uniform(-1, 1, (rows, k))
    bbuf = ctx.upload(b.reshape(-1))
    plain = spmm(ctx, a, bbuf, cols, c0, k, k, 1.25, H.ALG_AUTO, np.float64)
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(long_row_min=1024, staging=H.STAGING_STREAM))
    panel = spmm(ctx, a, bbuf, cols, c0, k, k, 1.25, H.ALG_PANEL, np.float64)
    auto = spmm(ctx, a, bbuf, cols, c0, k, k, 1.25, H.ALG_AUTO, np.float64)
    np.testing.assert_array_equal(auto.view(np.uint64), panel.view(np.uint64))
    assert not np.array_equal(auto.view(np.uint64), plain.view(np.uint64))  # long rows differ
    small = _spmm_raw(ctx, a, bbuf, rows, cols, c0, k, H.ALG_AUTO, None, 0)
    np.testing.assert_array_equal(small.view(np.uint64), plain.view(np.uint64))
    # An explicit ALG_PANEL needs its partial sums' workspace.
    with pytest.raises(H.AscHipError) as e:
        _spmm_raw(ctx, a, bbuf, rows, cols, c0, k, H.ALG_PANEL, None, 0)
    assert e.value.code == H.INVALID_ARGUMENT
    del plan
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(long_row_min=1, staging=H.STAGING_STREAM))
    every = spmm(ctx, a, bbuf, cols, c0, k, k, 1.25, H.ALG_AUTO, np.float64)
    np.testing.assert_array_equal(every.view(np.uint64), plain.view(np.uint64))
    del plan
    H.plan_clear(ctx)


def test_spmm_panel_nothing_listed_equals_plain(ctx):
    # An analysis listing no row (long_row_min above every row length) leaves
    # every row to the plain kernel.
    m = G.make(G.ZIPF, 3000, 3000, seed=4, values=G.VALUES_NORMAL, zipf_max=50)
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, (3000, 8))
    c0 = rng.uniform(-1, 1, (3000, 8))
    a = H.clone_compressed(ctx, H.CSR, 3000, 3000, m.offsets, m.indices, m.values, H.F64, 4)
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(long_row_min=100, staging=H.STAGING_STREAM))
    bbuf = ctx.upload(b.reshape(-1))
    got = spmm(ctx, a, bbuf, 3000, c0, 8, 8, 1.5, H.ALG_PANEL, np.float64)
    plain = spmm(ctx, a, bbuf, 3000, c0, 8, 8, 1.5, H.ALG_AUTO, np.float64)
    np.testing.assert_array_equal(got.view(np.uint64), plain.view(np.uint64))
    del plan


def test_spmm_panel_refusals(ctx):
    m = G.make(G.ZIPF, 2000, 2000, seed=5, values=G.VALUES_NORMAL, zipf_max=500)
    a = H.clone_compressed(ctx, H.CSR, 2000, 2000, m.offsets, m.indices, m.values, H.F64, 4)
    b = ctx.upload(np.zeros(2000 * 8))
    c = ctx.upload(np.zeros(2000 * 8))
    bd = H.dense(b, 2000, 8, 8, H.ROW_MAJOR, H.F64, writable=False)
    cd = H.dense(c, 2000, 8, 8, H.ROW_MAJOR, H.F64)
    # No analysis of this matrix.
    with pytest.raises(H.AscHipError) as e:
        H.standard_spmm(ctx, H.OP_N, 1.0, a, bd, cd, alg=H.ALG_PANEL)
    assert e.value.code == H.INVALID_ARGUMENT
    # An L2-staged analysis has no segment tables.
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_L2))
    with pytest.raises(H.AscHipError) as e:
        H.standard_spmm(ctx, H.OP_N, 1.0, a, bd, cd, alg=H.ALG_PANEL)
    assert e.value.code == H.INVALID_ARGUMENT
    del plan
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_STREAM))
    # op(A) = A^T, and a column-major block, are not the panel form's.
    with pytest.raises(H.AscHipError) as e:
        H.standard_spmm(ctx, H.OP_T, 1.0, a, bd, cd, alg=H.ALG_PANEL)
    assert e.value.code == H.UNSUPPORTED
    with pytest.raises(H.AscHipError) as e:
        H.standard_spmm(ctx, H.OP_N, 1.0, a, H.dense(b, 2000, 8, 2000, H.COLUMN_MAJOR, H.F64, writable=False),
                        cd, alg=H.ALG_PANEL)
    assert e.value.code == H.UNSUPPORTED
    del plan


def test_spmm_panel_b_beyond_32bit_offsets(ctx):
    # 2^24 columns of a 32-column FP64 B (4 GiB: 1024 panels of 16384
    # columns, the stream analysis's most), so the segment kernel takes its
    # 64-bit B offsets.  Rows of 2000-3000 entries spread over every panel; a
    # few short rows take the plain kernel.
    rows, cols, k = 512, 1 << 24, 32
    rng = np.random.default_rng(91)
    lens = np.where(np.arange(rows) % 8 == 0, rng.integers(1, 40, rows), rng.integers(2000, 3000, rows))
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    indices = np.concatenate([np.sort(rng.choice(cols, int(n), replace=False)) for n in lens])
    values = rng.standard_normal(int(offsets[-1]))
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, indices, values, H.F64, 4)
    b = rng.uniform(-1, 1, (cols, k))
    c0 = rng.uniform(-1, 1, (rows, k))
    print(f"spmm panel, B of {b.nbytes / 1e9:.1f} GB, nnz {offsets[-1]}", flush=True)
    bbuf = ctx.upload(b.reshape(-1))
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(long_row_min=1000, staging=H.STAGING_STREAM))
    got = spmm(ctx, a, bbuf, cols, c0, k, k, 0.5, H.ALG_PANEL, np.float64)
    ref = O.std_spmm(0, 0, rows, cols, offsets, indices.astype(np.int64), values, 0.5,
                     b.reshape(-1).copy(), 1, k, k, c0.reshape(-1).copy(), 1, k).reshape(rows, k)
    bound = plain_spmm_bound(H.OP_N, rows, cols, offsets, indices, values, 0.5, b, c0, EPS[np.float64])
    err = np.abs(got - ref)
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))
    del plan, bbuf
    H.plan_clear(ctx)


def test_spmm_panel_alpha_zero_leaves_c(ctx):
    # Standard SpMM with alpha == 0 is a no-op (blas.h:1730-1732): C keeps its
    # bits, NaN values in A included, in the panel form and under AUTO.
    m = G.make(G.ZIPF, 4096, 4096, seed=8, values=G.VALUES_NORMAL, zipf_max=4000)
    vals = m.values.copy()
    vals[::7] = np.nan
    a = H.clone_compressed(ctx, H.CSR, 4096, 4096, m.offsets, m.indices, vals, H.F64, 4)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (4096, 32))
    c0 = rng.uniform(-1, 1, (4096, 32))
    bbuf = ctx.upload(b.reshape(-1))
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(long_row_min=1024, staging=H.STAGING_STREAM))
    for alg in (H.ALG_PANEL, H.ALG_AUTO):
        got = spmm(ctx, a, bbuf, 4096, c0, 32, 32, 0.0, alg, np.float64)
        np.testing.assert_array_equal(got.view(np.uint64), c0.view(np.uint64))
    del plan
    H.plan_clear(ctx)


ORDERED_CASES = [
    # (rows, cols, zipf_max, long_row_min, panel_columns, k, ld, dtype, index_bytes)
    (1 << 15, 1 << 15, 10_000, 1024, 0, 32, 32, np.float64, 4),
    (1 << 15, 1 << 15, 10_000, 512, 2048, 20, 24, np.float64, 8),
    (5000, 70_000, 3000, 1000, 97, 16, 16, np.float32, 4),
    (5000, 70_000, 3000, 2000, 0, 4, 8, np.float64, 4),
]


@pytest.mark.parametrize("rows,cols,zmax,long_min,pcols,k,ld,dtype,ib", ORDERED_CASES)
def test_spmm_ordered_over_panels_bit_exact(ctx, rows, cols, zmax, long_min, pcols, k, ld, dtype, ib):
    # ORDERED on a stream-staged analysis listing long rows walks the panels in
    # order, continuing each C(i,j)'s serial chain: bit for bit the serial
    # reference (and the ordered form without the analysis), signed zeros and
    # the padding included.
    m = G.make(G.ZIPF, rows, cols, seed=rows % 997 + k, values=G.VALUES_NORMAL, zipf_max=zmax)
    val = m.values.astype(dtype)
    val[::11] = 0.0  # zero products: -0.0 + 0 must stay the serial result
    rng = np.random.default_rng(k + long_min)
    b = rng.uniform(-1, 1, (cols, ld)).astype(dtype)
    c0 = rng.uniform(-1, 1, (rows, ld)).astype(dtype)
    c0[::5, :] = -0.0
    a = H.clone_compressed(ctx, H.CSR, rows, cols, m.offsets, m.indices, val, KIND[dtype], ib)
    bbuf = ctx.upload(b.reshape(-1))
    before = spmm(ctx, a, bbuf, cols, c0, k, ld, -0.75, H.ALG_ORDERED, dtype)
    plan = H.csr_spmv_analyze(ctx, a, H.panel_options(panel_columns=pcols, long_row_min=long_min,
                                                      staging=H.STAGING_STREAM))
    print(f"ordered over panels rows {rows} cols {cols} long_min {long_min} panel {pcols} k {k} "
          f"{dtype.__name__} ib {ib}", flush=True)
    got = spmm(ctx, a, bbuf, cols, c0, k, ld, -0.75, H.ALG_ORDERED, dtype)
    ref = O.std_spmm(0, 0, rows, cols, m.offsets, m.indices.astype(np.int64), val, -0.75,
                     b.reshape(-1).copy(), 1, ld, k, c0.reshape(-1).copy(), 1, ld).reshape(rows, ld)
    bits = np.uint64 if dtype == np.float64 else np.uint32
    np.testing.assert_array_equal(got.view(bits), ref.view(bits))
    np.testing.assert_array_equal(got.view(bits), before.view(bits))
    del plan
    H.plan_clear(ctx)
