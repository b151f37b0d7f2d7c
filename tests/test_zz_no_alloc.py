"""The allocation-free compat SpMV (asc_sparse_hip_csr_spmv_noalloc) against
the existing entry point (asc_sparse_hip_csr_spmv with a caller workspace):
the same kernels, so the same bits, and both within the plain-mode bound of
the oracle.  Two calls in a row, the second on a smaller matrix (the shared
scratch keeps the first call's contents), for merge path and the
stream-staged plan, whose workspaces live in the scratch; and the refusal of
a call whose workspace exceeds the scratch."""
import ctypes as C

import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

EPS = 2.0 ** -53


def _problem(ctx, m, seed):
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    x = G.vector(m.cols, seed)
    yref = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, yref)
    bound = 4.0 * np.maximum(np.diff(m.offsets), 1) * EPS * O.abs_spmv(m.offsets, m.indices,
                                                                      m.values, x)
    return a, x, yref, bound


def _both(ctx, a, x, alg):
    xb, y1b, y2b = ctx.upload(x), ctx.alloc(8 * a.rows), ctx.alloc(8 * a.rows)
    ctx.fill(y2b, 0xFF)  # NaN pattern: every row must be written
    xd = H.vector(xb, a.cols, H.F64, writable=False)
    H.csr_spmv(ctx, 1.0, a, xd, 0.0, H.vector(y1b, a.rows, H.F64), alg=alg)
    H.csr_spmv_noalloc(ctx, 1.0, a, xd, 0.0, H.vector(y2b, a.rows, H.F64), alg=alg)
    return ctx.download(y1b, np.float64, a.rows), ctx.download(y2b, np.float64, a.rows)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [H.ALG_MERGE, H.ALG_PANEL])
def test_noalloc_matches_workspace_entry_point(ctx, alg):
    big = G.make(G.ZIPF, 1 << 18, 1 << 18, seed=61, values=G.VALUES_NORMAL)
    small = G.uniform16(4096)
    for m, seed in ((big, 62), (small, 63)):  # second call smaller
        a, x, yref, bound = _problem(ctx, m, seed)
        if alg == H.ALG_PANEL:
            H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_STREAM))
        print(f"allocation-free SpMV, {H.ALG_NAMES[alg]}, {m.rows} rows, {m.nnz} entries",
              flush=True)
        y1, y2 = _both(ctx, a, x, alg)
        np.testing.assert_array_equal(y2.view(np.uint64), y1.view(np.uint64))
        assert np.all(np.abs(y2 - yref) <= bound)
        H.plan_clear(ctx)


def test_noalloc_refuses_a_workspace_beyond_the_scratch(vctx):
    # Merge path over 2^31 + 2^31 rows and entries needs about 128 MiB of
    # tile starts: refused before anything else happens.
    t = 1 << 40  # far-apart fake addresses: a validation-only context never reads them
    a = H.SparseDesc(t, 2 * t, 3 * t, 1 << 31, 1 << 20, 1 << 31, H.CSR, H.F64, 4, 8, H.DEVICE, 1)
    x = H.VectorDesc(4 * t, 1 << 20, 1, H.F64, H.DEVICE, 0)
    y = H.VectorDesc(5 * t, 1 << 31, 1, H.F64, H.DEVICE, 1)
    opt = H.options(H.ALG_MERGE)
    st = vctx.lib.asc_sparse_hip_csr_spmv_noalloc(vctx.handle, 1.0, C.byref(a), C.byref(x), 0.0,
                                                  C.byref(y), C.byref(opt))
    assert st.code == H.UNSUPPORTED
    assert b"allocation-free scratch" in vctx.lib.asc_sparse_hip_last_message()
    # A call that fits is valid up to the missing device.
    a.rows, a.nnz, y.extent = 1000, 1000, 1000
    st = vctx.lib.asc_sparse_hip_csr_spmv_noalloc(vctx.handle, 1.0, C.byref(a), C.byref(x), 0.0,
                                                  C.byref(y), C.byref(opt))
    assert st.code == H.UNAVAILABLE
