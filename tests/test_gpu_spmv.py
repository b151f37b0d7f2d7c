"""GPU parity of the compat CSR SpMV (y = alpha*A*x + beta*y) through the C ABI.

Checks: the reference's golden vectors for every algorithm; bit-exactness
of the ordered mode against the oracle; the plain-mode bound
|y - y_ref| <= 4*len_i*eps*|alpha|*(|A||x|)_i + 4*eps*|beta*y0_i|; edge cases
(empty rows, beta == 0 with NaN in y, zero-column and zero-row matrices);
int32 vs int64 index identity and run-to-run determinism.
"""
import json
import os

import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
ALGS = [H.ALG_ORDERED, H.ALG_MERGE] + [H.ALG_VECTOR[g] for g in (1, 2, 4, 8, 16, 32, 64)]
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}
KIND = {np.float64: H.F64, np.float32: H.F32}


def arr(v, dtype=np.float64):
    return np.array([float(t) if t != "nan" else np.nan for t in v], dtype=dtype)


def run_compat(ctx, a, x_host, y_host, alpha, beta, alg, incx=1, incy=1, acc=H.ACC_PLAIN,
               xkind=None):
    kind = xkind if xkind is not None else a.desc.element_kind
    if kind in (H.BF16, H.F16):
        kind = H.F32
    dt = np.float64 if kind == H.F64 else np.float32
    xb = ctx.upload(np.asarray(x_host, dtype=dt))
    yb = ctx.upload(np.asarray(y_host, dtype=dt))
    nx = (len(x_host) + incx - 1) // incx if len(x_host) else 0
    ny = (len(y_host) + incy - 1) // incy if len(y_host) else 0
    xd = H.vector(xb if len(x_host) else 0, a.cols if nx else 0, kind, incx, writable=False)
    yd = H.vector(yb if len(y_host) else 0, a.rows if ny else 0, kind, incy)
    H.csr_spmv(ctx, alpha, a, xd, beta, yd, alg=alg, acc=acc)
    return ctx.download(yb, dt, len(y_host))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(GOLD["compat_spmv"])))
@pytest.mark.parametrize("alg", ALGS)
def test_golden_vectors(ctx, case, dtype, alg):
    g = GOLD["compat_spmv"][case]
    print(f"golden case {case} dtype {dtype.__name__} alg {H.ALG_NAMES[alg]}", flush=True)
    a = H.clone_compressed(ctx, H.CSR, g["rows"], g["cols"], np.array(g["offsets"]),
                           np.array(g["indices"], dtype=np.int64), arr(g["values"], dtype),
                           KIND[dtype])
    x = arr(g.get("x_storage", g.get("x")), dtype)
    y = arr(g["y_storage"], dtype)
    out = run_compat(ctx, a, x, y, g["alpha"], g["beta"], alg, g["incx"], g["incy"])
    if "expected_storage" in g:
        np.testing.assert_array_equal(out, arr(g["expected_storage"], dtype))
    else:
        np.testing.assert_array_equal(out[:: g["incy"]][: g["rows"]], arr(g["expected_rows"], dtype))


def test_order_f32_all_ordered(ctx):
    g = GOLD["compat_order_f32"]
    a = H.clone_compressed(ctx, H.CSR, 1, 3, np.array(g["offsets"]), np.array(g["indices"]),
                           arr(g["values"], np.float32), H.F32)
    out = run_compat(ctx, a, arr(g["x"], np.float32), arr(g["y0"], np.float32), 1.0, 0.0,
                     H.ALG_ORDERED)
    assert out[0] == np.float32(0.0)


def bound_check(y, yref, lens, absax, alpha, beta, y0, eps):
    bound = 4.0 * np.maximum(lens, 1) * eps * abs(alpha) * absax + 4.0 * eps * np.abs(beta * y0)
    err = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = np.nonzero(~(err <= bound))[0]
    assert bad.size == 0, f"{bad.size} rows exceed the bound, first {bad[:5]} err {err[bad[:5]]} bound {bound[bad[:5]]}"


def _matrices():
    return {
        "uniform16_4096": lambda: G.uniform16(4096),
        "zipf_2e16": lambda: G.make(G.ZIPF, 1 << 16, 1 << 16, seed=17, values=G.VALUES_NORMAL),
        "banded_3e4": lambda: G.make(G.BANDED, 30000, 30000, seed=5, values=G.VALUES_UNIFORM, band=40, fill=0.5),
        "stencil3d_30": lambda: G.make(G.STENCIL3D, 27000, 27000, seed=1, values=G.VALUES_LAPLACIAN, grid=30),
    }


@pytest.mark.parametrize("name", list(_matrices()))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_random_parity_all_algorithms(ctx, name, dtype):
    m = _matrices()[name]()
    vals = m.values.astype(dtype)
    x = G.vector(m.cols, 99).astype(dtype)
    y0 = G.vector(m.rows, 98).astype(dtype)
    lens = np.diff(m.offsets)
    absax = O.abs_spmv(m.offsets, m.indices, vals.astype(np.float64), x.astype(np.float64))
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, vals, KIND[dtype],
                           index_bytes=4)
    for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
        yref = y0.copy()
        O.compat_spmv(m.offsets, m.indices, vals, alpha, x, beta, yref)
        for alg in ALGS:
            print(f"{name} {dtype.__name__} alpha {alpha} beta {beta} alg {H.ALG_NAMES[alg]}", flush=True)
            y = run_compat(ctx, a, x, y0, alpha, beta, alg)
            if alg == H.ALG_ORDERED:
                np.testing.assert_array_equal(y.view(np.uint8), yref.view(np.uint8))
            else:
                bound_check(y, yref, lens, absax, alpha, beta, y0, EPS[dtype])


def test_int32_and_int64_indices_identical(ctx):
    m = G.make(G.ZIPF, 1 << 14, 1 << 14, seed=3, values=G.VALUES_NORMAL)
    x = G.vector(m.cols, 1)
    a32 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 8)
    y0 = np.zeros(m.rows)
    for alg in ALGS:
        r32 = run_compat(ctx, a32, x, y0, 1.0, 0.0, alg)
        r64 = run_compat(ctx, a64, x, y0, 1.0, 0.0, alg)
        np.testing.assert_array_equal(r32.view(np.uint64), r64.view(np.uint64))


def test_deterministic_across_runs(ctx):
    m = G.zipf_large(1 << 15, seed=77)
    x = G.vector(m.cols, 2)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    y0 = np.zeros(m.rows)
    for alg in ALGS:
        r1 = run_compat(ctx, a, x, y0, 1.0, 0.0, alg)
        r2 = run_compat(ctx, a, x, y0, 1.0, 0.0, alg)
        np.testing.assert_array_equal(r1.view(np.uint64), r2.view(np.uint64))


def test_empty_rows_and_beta_zero_nan(ctx):
    # Many empty rows (merge path consumes row-end items only) and NaN y.
    rng = np.random.default_rng(4)
    rows, cols = 5000, 300
    lens = np.where(rng.random(rows) < 0.7, 0, rng.integers(1, 40, rows))
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens]).astype(np.int64)
    vals = rng.standard_normal(off[-1])
    x = rng.standard_normal(cols)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, off, idx, vals, H.F64)
    y0 = np.full(rows, np.nan)
    yref = y0.copy()
    O.compat_spmv(off, idx, vals, -2.0, x, 0.0, yref)
    absax = O.abs_spmv(off, idx, vals, x)
    for alg in ALGS:
        y = run_compat(ctx, a, x, y0, -2.0, 0.0, alg)
        assert not np.isnan(y).any()
        if alg == H.ALG_ORDERED:
            np.testing.assert_array_equal(y.view(np.uint64), yref.view(np.uint64))
        else:
            bound_check(y, yref, lens, absax, -2.0, 0.0, np.zeros(rows), EPS[np.float64])


def test_long_rows_span_many_merge_tiles(ctx):
    # Rows far longer than a merge tile (2048 items) plus short neighbours.
    rng = np.random.default_rng(8)
    cols = 100000
    lens = np.array([1, 50000, 3, 0, 20000, 7000, 1, 1, 9000, 2])
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens]).astype(np.int64)
    vals = rng.standard_normal(off[-1])
    x = rng.standard_normal(cols)
    a = H.clone_compressed(ctx, H.CSR, len(lens), cols, off, idx, vals, H.F64)
    y0 = rng.standard_normal(len(lens))
    yref = y0.copy()
    O.compat_spmv(off, idx, vals, 0.5, x, 3.0, yref)
    absax = O.abs_spmv(off, idx, vals, x)
    for alg in ALGS:
        y = run_compat(ctx, a, x, y0, 0.5, 3.0, alg)
        if alg == H.ALG_ORDERED:
            np.testing.assert_array_equal(y.view(np.uint64), yref.view(np.uint64))
        else:
            bound_check(y, yref, lens, absax, 0.5, 3.0, y0, EPS[np.float64])


def test_zero_rows_is_noop(ctx):
    a = H.clone_compressed(ctx, H.CSR, 0, 4, np.zeros(1, np.int64), np.zeros(0, np.int64),
                           np.zeros(0), H.F64)
    out = run_compat(ctx, a, np.zeros(4), [], 1.0, 0.0, H.ALG_MERGE)
    assert out.size == 0


# ---- column-panel plan ------------------------------------------------------------

def _panel_cases():
    # (matrix, [(panel columns, long-row minimum)])
    return {
        "zipf_2e16": (lambda: G.make(G.ZIPF, 1 << 16, 1 << 16, seed=23, values=G.VALUES_NORMAL),
                      [(0, 0), (4099, 8), (1 << 13, 1), (300, 64)]),
        "uniform16_4096": (lambda: G.uniform16(4096), [(0, 0), (300, 16), (1000, 17)]),
        "banded_3e4": (lambda: G.make(G.BANDED, 30000, 30000, seed=6, values=G.VALUES_UNIFORM,
                                      band=40, fill=0.5), [(2048, 4), (37, 30)]),
    }


PANEL_KERNELS = [(H.ALG_AUTO, H.ALG_AUTO), (H.ALG_VECTOR[1], H.ALG_VECTOR[64]),
                 (H.ALG_VECTOR[8], H.ALG_VECTOR[1]), (H.ALG_VECTOR[64], H.ALG_VECTOR[8])]


@pytest.mark.parametrize("name", list(_panel_cases()))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_panel_plan_parity(ctx, name, dtype):
    make, shapes = _panel_cases()[name]
    m = make()
    vals = m.values.astype(dtype)
    x = G.vector(m.cols, 31).astype(dtype)
    y0 = G.vector(m.rows, 32).astype(dtype)
    y0[::97] = np.nan  # beta == 0 must not read y
    lens = np.diff(m.offsets)
    absax = O.abs_spmv(m.offsets, m.indices, vals.astype(np.float64), x.astype(np.float64))
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, vals, KIND[dtype],
                           index_bytes=4)
    xbytes = 8 if dtype == np.float64 else 4
    for width, long_min in shapes:
        # both stagings where a panel fits in LDS (160 KiB of x)
        fits = 0 < width and width * xbytes <= 160 << 10
        stagings = (H.STAGING_L2, H.STAGING_LDS) if fits else (H.STAGING_AUTO,)
        for (short_alg, long_alg), staging in [(k, st) for k in PANEL_KERNELS for st in stagings]:
            H.csr_spmv_analyze(ctx, a, H.panel_options(width, long_min, short_alg, long_alg, staging))
            assert H.plan_get(ctx, a) == H.ALG_PANEL
            for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
                y_in = y0 if beta == 0.0 else np.nan_to_num(y0)
                yref = y_in.copy()
                O.compat_spmv(m.offsets, m.indices, vals, alpha, x, beta, yref)
                print(f"panel {name} {dtype.__name__} width {width} long>={long_min} "
                      f"kernels {short_alg},{long_alg} staging {staging} alpha {alpha} beta {beta}",
                      flush=True)
                for alg in (H.ALG_PANEL, H.ALG_AUTO):
                    y = run_compat(ctx, a, x, y_in, alpha, beta, alg)
                    bound_check(y, yref, lens, absax, alpha, beta, np.nan_to_num(y_in), EPS[dtype])
    H.plan_clear(ctx)
    with pytest.raises(H.AscHipError):
        run_compat(ctx, a, x, y0, 1.0, 0.0, H.ALG_PANEL)


def test_panel_plan_strides_and_empty_rows(ctx):
    rng = np.random.default_rng(9)
    rows, cols = 7000, 5000
    lens = np.where(rng.random(rows) < 0.5, 0, rng.integers(1, 60, rows))
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.int64)
    vals = rng.standard_normal(offsets[-1])
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, idx, vals, H.F64, 8)
    H.csr_spmv_analyze(ctx, a, H.panel_options(333, 20, H.ALG_VECTOR[2], H.ALG_VECTOR[4]))
    x = rng.uniform(-1, 1, cols * 3)
    y0 = rng.uniform(-1, 1, rows * 2)
    xs = np.ascontiguousarray(x[::3])
    yref = y0[::2].copy()
    O.compat_spmv(offsets, idx, vals, -2.0, xs, 0.5, yref)
    absax = O.abs_spmv(offsets, idx, vals, xs)
    y = run_compat(ctx, a, x, y0, -2.0, 0.5, H.ALG_PANEL, incx=3, incy=2)
    bound_check(y[::2], yref, lens, absax, -2.0, 0.5, y0[::2], EPS[np.float64])
    np.testing.assert_array_equal(y[1::2], y0[1::2])  # holes untouched
    # empty rows with alpha < 0, beta == 0: y = alpha * (+0.0) = -0.0
    y = run_compat(ctx, a, xs, np.full(rows, np.nan), -1.0, 0.0, H.ALG_PANEL)
    empty = lens == 0
    assert np.all(np.signbit(y[empty])) and np.all(y[empty] == 0.0)
    H.plan_clear(ctx)


@pytest.mark.parametrize("staging", [H.STAGING_L2, H.STAGING_LDS])
def test_panel_plan_standard_spmv_and_precision(ctx, staging):
    m = G.make(G.ZIPF, 1 << 14, 1 << 14, seed=41, values=G.VALUES_NORMAL)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    H.csr_spmv_analyze(ctx, a, H.panel_options(1500, 32, staging=staging))
    print(f"panel standard/precision staging {staging}", flush=True)
    x = G.vector(m.cols, 5)
    y0 = G.vector(m.rows, 6)
    yref = y0.copy()
    O.compat_spmv(m.offsets, m.indices, m.values, 0.75, x, 1.0, yref)  # y + 0.75*A*x
    xb, yb = ctx.upload(x), ctx.upload(y0)
    H.standard_spmv(ctx, H.OP_N, 0.75, a, H.vector(xb, m.cols, H.F64, writable=False),
                    H.vector(yb, m.rows, H.F64), alg=H.ALG_PANEL)
    y = ctx.download(yb, np.float64, m.rows)
    absax = O.abs_spmv(m.offsets, m.indices, m.values, x)
    bound_check(y, yref, np.diff(m.offsets), absax, 0.75, 1.0, y0, EPS[np.float64])
    # BF16 storage, FP32 accumulation (plain and compensated) through the plan
    y64 = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, y64)
    for kind, tol in ((H.BF16, 1e-2), (H.F16, 1e-2), (H.F32, 1e-6)):
        low = H.convert_values(ctx, a, kind)
        H.csr_spmv_analyze(ctx, low, H.panel_options(1500, 32, staging=staging))
        for acc in (H.ACC_PLAIN, H.ACC_TWOSUM):
            y = run_compat(ctx, low, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL, acc=acc)
            err = np.max(np.abs(y - y64)) / np.max(absax)
            assert err <= tol, (kind, acc, err)
    H.plan_clear(ctx)


def test_lds_panels_negative_stride_and_dense_rows(ctx):
    """LDS staging with x read through a negative stride, on rows dense enough
    that AUTO staging picks LDS by itself (48+ entries per LDS-sized panel)."""
    rows, cols, per_row = 64, 1 << 20, 4000  # 56 LDS panels of 18725 columns: 71 per segment
    m = G.make(G.UNIFORM_FIXED, rows, cols, seed=44, values=G.VALUES_NORMAL, per_row=per_row)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    need_auto = H.csr_spmv_analysis_size(ctx, a)
    need_l2 = H.csr_spmv_analysis_size(ctx, a, H.panel_options(staging=H.STAGING_L2))
    assert need_auto > need_l2  # AUTO took the LDS-sized panels (more of them)
    H.csr_spmv_analyze(ctx, a)
    x = G.vector(cols * 2, 8)
    xs = np.ascontiguousarray(x[::-2])  # logical x for stride -2 from the far end
    y0 = G.vector(rows, 9)
    yref = y0.copy()
    O.compat_spmv(m.offsets, m.indices, m.values, -1.25, xs, 0.5, yref)
    absax = O.abs_spmv(m.offsets, m.indices, m.values, xs)
    xb, yb = ctx.upload(x), ctx.upload(y0)
    last = xb.ptr + 8 * (2 * cols - 1)  # element 0 of a stride -2 view sits at the far end
    print("LDS panels, negative stride x", flush=True)
    H.csr_spmv(ctx, -1.25, a, H.vector(last, cols, H.F64, stride=-2, writable=False), 0.5,
               H.vector(yb, rows, H.F64), alg=H.ALG_PANEL)
    y = ctx.download(yb, np.float64, rows)
    bound_check(y, yref, np.diff(m.offsets), absax, -1.25, 0.5, y0, EPS[np.float64])
    H.plan_clear(ctx)


# ---- column-panel plan, stream staging -----------------------------------------------

def _explicit(rows, cols, lens, seed):
    """Sorted distinct random columns per row, N(0,1) values."""
    rng = np.random.default_rng(seed)
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    idx = np.concatenate([np.sort(rng.choice(cols, int(n), replace=False)) if n < cols
                          else np.arange(cols) for n in lens]).astype(np.int64)
    return offsets, idx, rng.standard_normal(int(offsets[-1]))


def _stream_cases():
    # (offsets, indices, values, rows, cols) makers and [(panel columns, long-row minimum)]
    def gm(make):
        def f():
            m = make()
            return m.offsets, m.indices, m.values, m.rows, m.cols
        return f

    def dense_rows():
        # whole dense rows: every (panel, row) segment spans several slices
        rows, cols = 6, 1 << 18
        lens = np.array([cols, 5, 0, cols - 17, 300, cols // 2])
        o, i, v = _explicit(rows, cols, lens, 61)
        return o, i, v, rows, cols

    def half_empty():
        rng = np.random.default_rng(62)
        rows, cols = 20000, 90000
        lens = np.where(rng.random(rows) < 0.6, 0, rng.integers(1, 400, rows))
        o, i, v = _explicit(rows, cols, lens, 63)
        return o, i, v, rows, cols

    def one_per_row():
        # every entry its own segment: each full step finishes 512 segments,
        # more than a wave's ring holds, so the steps write them directly
        rng = np.random.default_rng(64)
        rows, cols = 60000, 8 * 16384
        idx = rng.integers(0, cols, rows).astype(np.int64)
        return np.arange(rows + 1, dtype=np.int64), idx, rng.standard_normal(rows), rows, cols

    return {
        "zipf_2e16": (gm(lambda: G.make(G.ZIPF, 1 << 16, 1 << 16, seed=23, values=G.VALUES_NORMAL)),
                      [(0, 0), (4099, 8), (1 << 13, 1), (300, 64)]),
        # width 4: 1024 panels, one-entry segments (a step finishes more
        # segments than a wave's ring holds and writes them directly)
        "uniform16_4096": (gm(lambda: G.uniform16(4096)), [(0, 0), (300, 16), (1000, 17), (4, 1)]),
        "banded_3e4": (gm(lambda: G.make(G.BANDED, 30000, 30000, seed=6, values=G.VALUES_UNIFORM,
                                         band=40, fill=0.5)), [(2048, 4), (37, 30)]),
        "dense_rows": (dense_rows, [(0, 0), (0, 1), (5000, 100)]),
        "half_empty": (half_empty, [(0, 1), (1 << 12, 100), (777, 1)]),
        "one_per_row": (one_per_row, [(0, 0), (5000, 1)]),
    }


@pytest.mark.parametrize("name", list(_stream_cases()))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stream_plan_parity(ctx, name, dtype):
    make, shapes = _stream_cases()[name]
    offsets, idx, values, rows, cols = make()
    vals = values.astype(dtype)
    x = G.vector(cols, 71).astype(dtype)
    y0 = G.vector(rows, 72).astype(dtype)
    y0[::89] = np.nan  # beta == 0 must not read y
    lens = np.diff(offsets)
    absax = O.abs_spmv(offsets, idx, vals.astype(np.float64), x.astype(np.float64))
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, idx, vals, KIND[dtype], index_bytes=4)
    for width, long_min in shapes:
        opts = H.panel_options(width, long_min, staging=H.STAGING_STREAM)
        H.csr_spmv_analyze(ctx, a, opts)
        assert H.plan_get(ctx, a) == H.ALG_PANEL
        for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
            y_in = y0 if beta == 0.0 else np.nan_to_num(y0)
            yref = y_in.copy()
            O.compat_spmv(offsets, idx, vals, alpha, x, beta, yref)
            print(f"stream {name} {dtype.__name__} width {width} long>={long_min} "
                  f"alpha {alpha} beta {beta}", flush=True)
            first = None
            for alg in (H.ALG_PANEL, H.ALG_AUTO, H.ALG_PANEL):
                # AUTO never takes the stream plan (merge-path on the current
                # values); both ALG_PANEL calls give identical bits.
                y = run_compat(ctx, a, x, y_in, alpha, beta, alg)
                bound_check(y, yref, lens, absax, alpha, beta, np.nan_to_num(y_in), EPS[dtype])
                if alg == H.ALG_AUTO:
                    continue
                if first is None:
                    first = y
                else:  # fixed decomposition: identical bits every call
                    np.testing.assert_array_equal(y.view(np.uint8), first.view(np.uint8))
    H.plan_clear(ctx)


@pytest.mark.parametrize("groups", [300, 600, 1100, 2100, 4200])
def test_stream_plan_combine_shapes(ctx, groups):
    # The combine gives a work-group 16 / w adjacent groups of 1024 listed
    # rows, w waves splitting each group's panels, with w the fewest that
    # still yield a work-group per CU: 300 groups run one group per
    # work-group, 600 / 1100 / 2100 / 4200 run two, four, eight and sixteen
    # (the 2^22-row Zipf matrix, 4096 groups, runs sixteen).
    rng = np.random.default_rng(groups)
    rows, cols = groups * 1024, 1 << 17
    r = np.repeat(np.arange(rows, dtype=np.int64), rng.integers(1, 5, rows))
    key = np.unique(r * cols + rng.integers(0, cols, r.size))  # sorted, distinct per row
    idx = key % cols
    offsets = np.concatenate([[0], np.cumsum(np.bincount(key // cols, minlength=rows))]).astype(np.int64)
    vals = rng.standard_normal(idx.size)
    x = G.vector(cols, 73)
    y0 = G.vector(rows, 74)
    lens = np.diff(offsets)
    absax = O.abs_spmv(offsets, idx, vals, x)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, idx, vals, H.F64, index_bytes=4)
    H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_STREAM))
    assert H.plan_get(ctx, a) == H.ALG_PANEL
    for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
        yref = y0.copy()
        O.compat_spmv(offsets, idx, vals, alpha, x, beta, yref)
        print(f"stream combine shape: {groups} groups alpha {alpha} beta {beta}", flush=True)
        y = run_compat(ctx, a, x, y0, alpha, beta, H.ALG_PANEL)
        bound_check(y, yref, lens, absax, alpha, beta, y0, EPS[np.float64])
        again = run_compat(ctx, a, x, y0, alpha, beta, H.ALG_PANEL)
        np.testing.assert_array_equal(again.view(np.uint8), y.view(np.uint8))
    H.plan_clear(ctx)


def test_stream_plan_strides_standard_and_empty_rows(ctx):
    rng = np.random.default_rng(19)
    rows, cols = 7000, 50000
    lens = np.where(rng.random(rows) < 0.5, 0, rng.integers(1, 600, rows))
    offsets, idx, vals = _explicit(rows, cols, lens, 20)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, idx, vals, H.F64, 8)
    H.csr_spmv_analyze(ctx, a, H.panel_options(3333, 20, staging=H.STAGING_STREAM))
    x = rng.uniform(-1, 1, cols * 3)
    y0 = rng.uniform(-1, 1, rows * 2)
    xs = np.ascontiguousarray(x[::3])
    yref = y0[::2].copy()
    O.compat_spmv(offsets, idx, vals, -2.0, xs, 0.5, yref)
    absax = O.abs_spmv(offsets, idx, vals, xs)
    print("stream plan, strided x and y", flush=True)
    y = run_compat(ctx, a, x, y0, -2.0, 0.5, H.ALG_PANEL, incx=3, incy=2)
    bound_check(y[::2], yref, lens, absax, -2.0, 0.5, y0[::2], EPS[np.float64])
    np.testing.assert_array_equal(y[1::2], y0[1::2])  # holes untouched
    # empty rows with alpha < 0, beta == 0: y = alpha * (+0.0) = -0.0
    y = run_compat(ctx, a, xs, np.full(rows, np.nan), -1.0, 0.0, H.ALG_PANEL)
    empty = lens == 0
    assert np.all(np.signbit(y[empty])) and np.all(y[empty] == 0.0)
    # negative stride x: element 0 at the far end
    xb, yb = ctx.upload(x), ctx.upload(np.zeros(rows))
    xneg = np.ascontiguousarray(x[::-3][:cols])
    last = xb.ptr + 8 * (3 * cols - 1)
    print("stream plan, negative stride x", flush=True)
    H.csr_spmv(ctx, 1.0, a, H.vector(last, cols, H.F64, stride=-3, writable=False), 0.0,
               H.vector(yb, rows, H.F64), alg=H.ALG_PANEL)
    yref = np.zeros(rows)
    O.compat_spmv(offsets, idx, vals, 1.0, xneg, 0.0, yref)
    bound_check(ctx.download(yb, np.float64, rows), yref, lens,
                O.abs_spmv(offsets, idx, vals, xneg), 1.0, 0.0, np.zeros(rows), EPS[np.float64])
    # standardized SpMV through the same plan: y + 0.75*A*x
    y1 = rng.uniform(-1, 1, rows)
    yref = y1.copy()
    O.compat_spmv(offsets, idx, vals, 0.75, xs, 1.0, yref)
    xb, yb = ctx.upload(xs), ctx.upload(y1)
    print("stream plan, standardized SpMV", flush=True)
    H.standard_spmv(ctx, H.OP_N, 0.75, a, H.vector(xb, cols, H.F64, writable=False),
                    H.vector(yb, rows, H.F64), alg=H.ALG_PANEL)
    bound_check(ctx.download(yb, np.float64, rows), yref, lens, absax, 0.75, 1.0, y1,
                EPS[np.float64])
    H.plan_clear(ctx)


def test_stream_plan_reduced_precision_and_refusals(ctx):
    m = G.make(G.ZIPF, 1 << 14, 1 << 14, seed=43, values=G.VALUES_NORMAL)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    x = G.vector(m.cols, 5)
    y64 = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, m.values, 1.0, x, 0.0, y64)
    absax = O.abs_spmv(m.offsets, m.indices, m.values, x)
    for kind, tol in ((H.BF16, 1e-2), (H.F16, 1e-2), (H.F32, 1e-6)):
        low = H.convert_values(ctx, a, kind)
        H.csr_spmv_analyze(ctx, low, H.panel_options(1500, 32, staging=H.STAGING_STREAM))
        print(f"stream plan, storage kind {kind}", flush=True)
        y = run_compat(ctx, low, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL)
        err = np.max(np.abs(y - y64)) / np.max(absax)
        assert err <= tol, (kind, err)
        with pytest.raises(H.AscHipError):  # plain accumulation only
            run_compat(ctx, low, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL, acc=H.ACC_TWOSUM)
    # values rebound after the analysis: explicit PANEL is refused, AUTO runs
    # merge-path on the new values, a new analysis serves them
    H.csr_spmv_analyze(ctx, a, H.panel_options(1500, 32, staging=H.STAGING_STREAM))
    new_vals = np.random.default_rng(4).standard_normal(m.nnz)
    b = H.rebind_values(a, ctx.upload(new_vals))
    yref = np.zeros(m.rows)
    O.compat_spmv(m.offsets, m.indices, new_vals, 1.0, x, 0.0, yref)
    absnew = O.abs_spmv(m.offsets, m.indices, new_vals, x)
    lens = np.diff(m.offsets)
    with pytest.raises(H.AscHipError):
        run_compat(ctx, b, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL)
    print("stream plan, rebound values through AUTO", flush=True)
    y = run_compat(ctx, b, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_AUTO)
    bound_check(y, yref, lens, absnew, 1.0, 0.0, np.zeros(m.rows), EPS[np.float64])
    # a values-only refresh takes the rebound values into the copy
    H.csr_spmv_refresh_values(ctx, b)
    print("stream plan, refreshed after a rebind", flush=True)
    y = run_compat(ctx, b, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL)
    bound_check(y, yref, lens, absnew, 1.0, 0.0, np.zeros(m.rows), EPS[np.float64])
    with pytest.raises(H.AscHipError):  # the copy now follows b's values, not a's
        run_compat(ctx, a, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL)
    H.csr_spmv_analyze(ctx, b, H.panel_options(1500, 32, staging=H.STAGING_STREAM))
    y = run_compat(ctx, b, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL)
    bound_check(y, yref, lens, absnew, 1.0, 0.0, np.zeros(m.rows), EPS[np.float64])
    # the size query counts the listed rows: far below the every-row bound
    need = H.csr_spmv_analysis_size(ctx, a, H.panel_options(0, 64, staging=H.STAGING_STREAM))
    vctx = H.Context(-1)
    bound = H.csr_spmv_analysis_size(vctx, a, H.panel_options(0, 64, staging=H.STAGING_STREAM))
    vctx.close()
    assert need < bound
    H.plan_clear(ctx)


@pytest.mark.parametrize("vkind", [H.F64, H.BF16])
def test_stream_plan_in_place_value_update(ctx, vkind):
    """Values changed in place after a stream-staged analysis: AUTO (the
    reference's read-the-current-values semantics, reference_blas.cc:23-33)
    gives the new product at once; the explicit plan runs its copy until
    asc_sparse_hip_csr_spmv_refresh_values, then gives the new product."""
    m = G.make(G.ZIPF, 1 << 15, 1 << 16, seed=47, values=G.VALUES_NORMAL)
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    a = a64 if vkind == H.F64 else H.convert_values(ctx, a64, vkind)
    x = G.vector(m.cols, 6)
    lens = np.diff(m.offsets)
    H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_STREAM))
    new_vals = np.random.default_rng(48).standard_normal(m.nnz)
    eps, tol = (EPS[np.float64], None) if vkind == H.F64 else (None, 1e-2)

    def check(y, vals):
        yref = np.zeros(m.rows)
        O.compat_spmv(m.offsets, m.indices, vals, 1.0, x, 0.0, yref)
        absax = O.abs_spmv(m.offsets, m.indices, vals, x)
        if tol is None:
            bound_check(y, yref, lens, absax, 1.0, 0.0, np.zeros(m.rows), eps)
        else:
            assert np.max(np.abs(y - yref)) / np.max(absax) <= tol

    old = m.values
    check(run_compat(ctx, a, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL), old)
    # overwrite the values in place (same pointer)
    if vkind == H.F64:
        ctx.copy(a.desc.values, new_vals.ctypes.data, new_vals.nbytes)
    else:
        nb = ctx.upload(new_vals)
        fresh = H.convert_values(ctx, H.rebind_values(a64, nb), vkind)
        ctx.copy(a.desc.values, fresh.desc.values, 2 * m.nnz)
    ctx.synchronize()
    print(f"in-place update, kind {vkind}: AUTO reads the current values", flush=True)
    check(run_compat(ctx, a, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_AUTO), new_vals)
    print(f"in-place update, kind {vkind}: explicit plan until refreshed", flush=True)
    check(run_compat(ctx, a, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL), old)
    H.csr_spmv_refresh_values(ctx, a)
    check(run_compat(ctx, a, x, np.zeros(m.rows), 1.0, 0.0, H.ALG_PANEL), new_vals)
    H.plan_clear(ctx)
    with pytest.raises(H.AscHipError) as e:  # nothing to refresh without a stream plan
        H.csr_spmv_refresh_values(ctx, a)
    assert e.value.code == H.INVALID_STATE


def test_stream_plan_bounded_wait_reports_internal(ctx):
    """A stream-staged SpMV whose inter-wave wait gives up (forced from the
    host: no writer wave, waits bounded at zero, on a matrix whose rings
    fill within a slice) ends normally and the next synchronization fails
    with ASC_HIP_INTERNAL; the following SpMV is correct again.  The same
    layout gives up in every work-group of the CPU model
    (tests/test_stream_protocol.py)."""
    rng = np.random.default_rng(12)
    rows, cols = 40000, 16384
    idx = np.sort(np.stack([rng.choice(cols, 5, replace=False) for _ in range(rows)]), 1).reshape(-1)
    offsets = np.arange(0, 5 * rows + 1, 5, dtype=np.int64)
    vals = rng.standard_normal(idx.size)
    x = G.vector(cols, 8)
    a = H.clone_compressed(ctx, H.CSR, rows, cols, offsets, idx, vals, H.F64, 4)
    H.csr_spmv_analyze(ctx, a, H.panel_options(staging=H.STAGING_STREAM))
    xb, yb = ctx.upload(x), ctx.alloc(8 * rows)
    xd, yd = H.vector(xb, cols, H.F64, writable=False), H.vector(yb, rows, H.F64)
    ctx.synchronize()
    print("stream plan with a forced stall", flush=True)
    H.test_inject(H.INJECT_STREAM_STALL, 0)
    try:
        H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_PANEL)
        with pytest.raises(H.AscHipError) as e:
            ctx.synchronize()
        assert e.value.code == H.INTERNAL
    finally:
        H.test_inject(H.INJECT_STREAM_STALL, -1)
    ctx.synchronize()  # reported once
    print("stream plan after the stall", flush=True)
    H.csr_spmv(ctx, 1.0, a, xd, 0.0, yd, alg=H.ALG_PANEL)
    ctx.synchronize()
    yref = np.zeros(rows)
    O.compat_spmv(offsets, idx, vals, 1.0, x, 0.0, yref)
    bound_check(ctx.download(yb, np.float64, rows), yref, np.diff(offsets),
                O.abs_spmv(offsets, idx, vals, x), 1.0, 0.0, np.zeros(rows), EPS[np.float64])
    H.plan_clear(ctx)
