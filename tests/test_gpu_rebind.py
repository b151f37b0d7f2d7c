"""Values-only refresh on trusted device structure (RebindValues,
compressed.h:472-516; sparse.md:455-466) through the C ABI: a rebound clone
gives the oracle's results with the new values on the compat and
standardized paths, keeps its column-panel plan (the plan depends on the
structure only), and an operand whose span runs past its allocation is
refused before anything is enqueued.
"""
import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu


def zipf(n, seed):
    return G.make(G.ZIPF, n, n, seed=seed, values=G.VALUES_NORMAL)


def test_rebind_compat_and_panel(ctx):
    m = zipf(1 << 13, 71)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    H.csr_spmv_analyze(ctx, a, H.panel_options(700, 16))
    new_vals = np.random.default_rng(3).standard_normal(m.nnz)
    vb = ctx.upload(new_vals)
    b = H.rebind_values(a, vb)
    assert b.desc.trusted == 1 and b.desc.values == vb.ptr and b.desc.outer_offsets == a.desc.outer_offsets
    assert H.plan_get(ctx, b) == H.ALG_PANEL  # same structure, same plan
    x = G.vector(m.cols, 8)
    lens = np.diff(m.offsets)
    yref = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, new_vals, 1.5, x, 0.0, yref)
    absax = O.abs_spmv(m.offsets, m.indices, new_vals, x)
    for alg in (H.ALG_ORDERED, H.ALG_PANEL, H.ALG_MERGE, H.ALG_AUTO):
        print(f"rebound compat alg {H.ALG_NAMES.get(alg, alg)}", flush=True)
        xb, yb = ctx.upload(x), ctx.upload(np.full(m.rows, np.nan))
        H.csr_spmv(ctx, 1.5, b, H.vector(xb, m.cols, H.F64, writable=False), 0.0,
                   H.vector(yb, m.rows, H.F64), alg=alg)
        y = ctx.download(yb, np.float64, m.rows)
        if alg == H.ALG_ORDERED:
            np.testing.assert_array_equal(y.view(np.uint64), yref.view(np.uint64))
        else:
            bound = 4.0 * np.maximum(lens, 1) * 2.0 ** -53 * 1.5 * absax
            assert np.all(np.abs(y - yref) <= bound)
    # the original clone still computes with its own values
    xb, yb = ctx.upload(x), ctx.upload(np.zeros(m.rows))
    H.csr_spmv(ctx, 1.0, a, H.vector(xb, m.cols, H.F64, writable=False), 0.0,
               H.vector(yb, m.rows, H.F64), alg=H.ALG_ORDERED)
    yorig = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, yorig)
    np.testing.assert_array_equal(ctx.download(yb, np.float64, m.rows), yorig)
    H.plan_clear(ctx)


@pytest.mark.parametrize("op", [H.OP_N, H.OP_T])
def test_rebind_standard(ctx, op):
    m = zipf(3000, 72)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), m.values, H.F64)
    new_vals = np.random.default_rng(4).uniform(-1, 1, m.nnz)
    b = H.rebind_values(a, ctx.upload(new_vals))
    x = G.vector(m.cols, 9)
    y0 = G.vector(m.rows, 10)
    ref = y0.copy()
    O.std_spmv(0, op, m.rows, m.cols, m.offsets, m.indices, new_vals, -0.5, x, ref)
    xb, yb = ctx.upload(x), ctx.upload(y0)
    print(f"rebound standard op {op}", flush=True)
    H.standard_spmv(ctx, op, -0.5, b, H.vector(xb, m.cols, H.F64, writable=False),
                    H.vector(yb, m.rows, H.F64), alg=H.ALG_ORDERED)
    np.testing.assert_array_equal(ctx.download(yb, np.float64, m.rows).view(np.uint64),
                                  ref.view(np.uint64))


BLOCK = 1 << 22  # a 4 MiB allocation: the runtime reports its extent as requested


def at_end(ctx, span_bytes, overrun=0):
    """(buffer, address) with a span of span_bytes ending `overrun` bytes past
    the end of a BLOCK-byte allocation."""
    buf = ctx.alloc(BLOCK)
    return buf, buf.ptr + BLOCK - span_bytes + overrun


def test_span_past_allocation_is_refused(ctx):
    m = zipf(2000, 73)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices.astype(np.int64), m.values, H.F64)
    vbuf, vptr = at_end(ctx, 8 * m.nnz, overrun=8)
    b = H.rebind_values(a, vptr)  # the host cannot see the allocation's size
    x = G.vector(m.cols, 11)
    y0 = G.vector(m.rows, 12)
    xb, yb = ctx.upload(x), ctx.upload(y0)
    xd = H.vector(xb, m.cols, H.F64, writable=False)
    print("values span past its allocation", flush=True)
    with pytest.raises(H.AscHipError) as e:
        H.csr_spmv(ctx, 1.0, b, xd, 0.0, H.vector(yb, m.rows, H.F64), alg=H.ALG_ORDERED)
    assert e.value.code == H.MEMORY_ACCESS
    np.testing.assert_array_equal(ctx.download(yb, np.float64, m.rows), y0)
    # a strided y whose last element lies one element past its allocation
    print("strided y past its allocation", flush=True)
    ybuf, yptr = at_end(ctx, 8 * (2 * m.rows - 1), overrun=8)
    ctx.copy(yptr, y0.ctypes.data, 8)
    ctx.synchronize()
    with pytest.raises(H.AscHipError) as e:
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, H.vector(yptr, m.rows, H.F64, stride=2), alg=H.ALG_ORDERED)
    assert e.value.code == H.MEMORY_ACCESS
    assert ctx.download(ybuf, np.float64, 1, BLOCK - 8 * (2 * m.rows - 1) + 8)[0] == y0[0]
    # a negative stride whose lowest element lies before its allocation
    ybuf2 = ctx.alloc(BLOCK)
    with pytest.raises(H.AscHipError) as e:
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, H.vector(ybuf2.ptr + 8 * (m.rows - 2), m.rows, H.F64, stride=-1),
                   alg=H.ALG_ORDERED)
    assert e.value.code == H.MEMORY_ACCESS
    # x read one element past the end of its allocation
    xbuf, xptr = at_end(ctx, 8 * m.cols, overrun=8)
    with pytest.raises(H.AscHipError) as e:
        H.csr_spmv(ctx, 1.0, a, H.vector(xptr, m.cols, H.F64, writable=False), 0.0,
                   H.vector(yb, m.rows, H.F64), alg=H.ALG_ORDERED)
    assert e.value.code == H.MEMORY_ACCESS
    np.testing.assert_array_equal(ctx.download(yb, np.float64, m.rows), y0)
    # values ending exactly at the end of their allocation are accepted
    gbuf, gptr = at_end(ctx, 8 * m.nnz)
    ones = np.ones(m.nnz)
    ctx.copy(gptr, ones.ctypes.data, ones.nbytes)
    ctx.synchronize()
    good = H.rebind_values(a, gptr)
    H.csr_spmv(ctx, 1.0, good, xd, 0.0, H.vector(yb, m.rows, H.F64), alg=H.ALG_ORDERED)
    ref = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, np.ones(m.nnz), 1.0, x, 0.0, ref)
    np.testing.assert_array_equal(ctx.download(yb, np.float64, m.rows), ref)


def test_rebind_needs_new_triangular_analysis(ctx):
    # lower [[2, 0], [1, 3]]: the analysis checked these values, so a rebound
    # matrix is refused until analysed again
    off, idx = np.array([0, 1, 3]), np.array([0, 0, 1])
    a = H.clone_compressed(ctx, H.CSR, 2, 2, off, idx, np.array([2.0, 1.0, 3.0]), H.F64)
    plan = H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, a)
    new_vals = np.array([4.0, -1.0, 0.5])
    b = H.rebind_values(a, ctx.upload(new_vals))
    rhs = ctx.upload(np.array([8.0, 1.0]))
    with pytest.raises(H.AscHipError) as e:
        H.triangular_solve(ctx, H.OP_N, 1.0, b, H.dense(rhs, 2, 1, 1, H.ROW_MAJOR, H.F64), plan)
    assert e.value.code == H.INVALID_ARGUMENT
    plan_b = H.triangular_analyze(ctx, H.OP_N, H.LOWER, H.NON_UNIT, b)
    H.triangular_solve(ctx, H.OP_N, 1.0, b, H.dense(rhs, 2, 1, 1, H.ROW_MAJOR, H.F64), plan_b)
    ref = O.trsv(0, 0, True, False, 2, off, idx, new_vals, 1.0, np.array([8.0, 1.0]))
    np.testing.assert_array_equal(ctx.download(rhs, np.float64, 2), ref)
    assert list(ref) == [2.0, 6.0]  # 4*2 = 8; -1*2 + 0.5*6 = 1
