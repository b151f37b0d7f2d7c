"""Parity at the benchmark workloads' full sizes, through the C ABI.

- Single large matrix: n = 2^22 square, Zipf(1.5) row lengths clipped to
  [1, 10^4] (640,593,960 entries), int32 columns, N(0,1) values.  Every row
  of every SpMV is checked against the oracle (run over row ranges on host
  threads): the ordered kernel bit for bit, the stream-staged column-panel
  plan, merge path and the L2-staged plan within the plain-mode bound
  |y - y_ref| <= 4*len_i*eps*|alpha|*(|A||x|)_i + 4*eps*|beta*y0_i|, in FP64
  and FP32.  The ordered SpMM with a 2^22 x 32 row-major block is checked bit
  for bit on every entry, the plain and column-panel forms within the plain
  bound on every entry.  CSR->CSC is checked through properties that fix the
  canonical transpose (per-column counts, rows strictly increasing within each
  column, per-column checksums of row ids and value bits, sampled entries
  found with their exact bits) and transposing back gives the original bit
  for bit.  CSR->COO->CSR gives the original bit for bit, and so does a
  seeded shuffle of the same entries through the device finalize (the sort
  path of COO->CSR).
- Batched: all 10,000 matrices (rows log-uniform in [64, 65536]; 2D/3D
  Laplacians and random banded), FP64 and FP32, tuned, captured in a
  hipGraph and replayed; every row of every matrix within the plain bound.
- Precision: n = 2^20, 32 entries per row, lognormal values; FP32/BF16/FP16
  storage with FP32 accumulation against the FP64 oracle.
"""
import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu

EPS64, EPS32 = 2.0 ** -53, 2.0 ** -24


def say(msg):
    print(msg, flush=True)


def bound_check(y, yref, lens, absax, alpha, beta, y0, eps, what):
    bound = 4.0 * np.maximum(lens, 1) * eps * abs(alpha) * absax
    if beta != 0.0:
        bound += 4.0 * eps * np.abs(beta * y0)
    err = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = np.nonzero(~(err <= bound))[0]
    assert bad.size == 0, (f"{what}: {bad.size} rows exceed the bound, first {bad[:5]} "
                           f"err {err[bad[:5]]} bound {bound[bad[:5]]}")


@pytest.fixture(scope="module")
def cfg2(ctx):
    say("single large matrix: generating n=2^22 Zipf(1.5)")
    m = G.zipf_large(1 << 22)
    say(f"single large matrix: {m.nnz} entries; cloning to the device")
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    x = G.vector(m.cols, 12345)
    yield m, a, x
    H.plan_clear(ctx)
    ctx.synchronize()


def _spmv(ctx, a, xb, yb, y0, alpha, beta, alg, kind):
    dt = np.float64 if kind == H.F64 else np.float32
    ctx.copy(yb.ptr, y0.ctypes.data, y0.nbytes)
    H.csr_spmv(ctx, alpha, a, H.vector(xb, a.cols, kind, writable=False), beta,
               H.vector(yb, a.rows, kind), alg=alg)
    return ctx.download(yb, dt, a.rows)


def test_cfg2_spmv_fp64_every_row(ctx, cfg2):
    m, a, x = cfg2
    lens = np.diff(m.offsets)
    absax = O.abs_spmv_rows(m.offsets, m.indices, m.values, x)
    xb, yb = ctx.upload(x), ctx.alloc(8 * m.rows)
    rng = np.random.default_rng(31)
    for alpha, beta in ((1.0, 0.0), (-1.5, 0.25)):
        y0 = G.vector(m.rows, 99)
        if beta == 0.0:
            y0[rng.integers(0, m.rows, 4096)] = np.nan  # beta == 0 never reads y
        yref = y0.copy()
        O.compat_spmv_rows(m.offsets, m.indices, m.values, alpha, x, beta, yref)
        say(f"FP64 alpha {alpha} beta {beta}: ordered kernel, bit for bit")
        y = _spmv(ctx, a, xb, yb, y0, alpha, beta, H.ALG_ORDERED, H.F64)
        np.testing.assert_array_equal(y.view(np.uint64), yref.view(np.uint64))
        say(f"FP64 alpha {alpha} beta {beta}: merge path")
        y = _spmv(ctx, a, xb, yb, y0, alpha, beta, H.ALG_MERGE, H.F64)
        bound_check(y, yref, lens, absax, alpha, beta, y0, EPS64, "merge")
        for staging, name in ((H.STAGING_STREAM, "stream"), (H.STAGING_L2, "L2")):
            H.csr_spmv_analyze(ctx, a, H.panel_options(staging=staging))
            say(f"FP64 alpha {alpha} beta {beta}: column-panel plan, {name} staging")
            y = _spmv(ctx, a, xb, yb, y0, alpha, beta, H.ALG_PANEL, H.F64)
            bound_check(y, yref, lens, absax, alpha, beta, y0, EPS64, name)
            again = _spmv(ctx, a, xb, yb, y0, alpha, beta, H.ALG_PANEL, H.F64)
            np.testing.assert_array_equal(again.view(np.uint64), y.view(np.uint64))
        a.analysis.free()
        a.analysis = None
        H.plan_clear(ctx)


def test_cfg2_spmv_fp32_every_row(ctx, cfg2):
    m, a, x = cfg2
    a32 = H.convert_values(ctx, a, H.F32)
    v32, x32 = m.values.astype(np.float32), x.astype(np.float32)
    lens = np.diff(m.offsets)
    absax = O.abs_spmv_rows(m.offsets, m.indices, v32.astype(np.float64), x32.astype(np.float64))
    xb, yb = ctx.upload(x32), ctx.alloc(4 * m.rows)
    y0 = np.full(m.rows, np.nan, dtype=np.float32)
    yref = np.zeros(m.rows, dtype=np.float32)
    O.compat_spmv_rows(m.offsets, m.indices, v32, 1.0, x32, 0.0, yref)
    say("FP32: ordered kernel, bit for bit")
    y = _spmv(ctx, a32, xb, yb, y0, 1.0, 0.0, H.ALG_ORDERED, H.F32)
    np.testing.assert_array_equal(y.view(np.uint32), yref.view(np.uint32))
    H.csr_spmv_analyze(ctx, a32, H.panel_options(staging=H.STAGING_STREAM))
    say("FP32: column-panel plan, stream staging")
    y = _spmv(ctx, a32, xb, yb, y0, 1.0, 0.0, H.ALG_PANEL, H.F32)
    bound_check(y, yref, lens, absax, 1.0, 0.0, y0, EPS32, "FP32 stream")
    a32.analysis.free()
    a32.analysis = None
    H.plan_clear(ctx)


@pytest.fixture(scope="module")
def cfg2_spmm(cfg2):
    """B (2^22 x 32, row-major), C0 and the serial result C0 + 0.5*A*B of the
    oracle over row ranges."""
    m, _, _ = cfg2
    k = 32
    b = G.vector(m.cols * k, 777).reshape(m.cols, k)
    c0 = G.vector(m.rows * k, 778).reshape(m.rows, k)
    say("SpMM k=32 row-major: oracle over row ranges")
    ref = c0.copy()

    def part(r0, r1):
        O.std_spmm(0, 0, r1 - r0, m.cols, m.offsets[r0:r1 + 1], m.indices, m.values, 0.5, b, 1, k, k,
                   ref[r0:r1], 1, k)
    O.parallel_rows(m.offsets, part)
    return b, c0, ref


def _device_spmm(ctx, m, a, b, c0, alg):
    k = b.shape[1]
    bb, cb = ctx.upload(b), ctx.upload(c0)
    H.standard_spmm(ctx, H.OP_N, 0.5, a, H.dense(bb, m.cols, k, k, H.ROW_MAJOR, H.F64, writable=False),
                    H.dense(cb, m.rows, k, k, H.ROW_MAJOR, H.F64), alg=alg)
    got = ctx.download(cb, np.float64, m.rows * k).reshape(m.rows, k)
    bb.free()
    cb.free()
    return got


def test_cfg2_spmm_k32_ordered_every_entry(ctx, cfg2, cfg2_spmm):
    m, a, _ = cfg2
    b, c0, ref = cfg2_spmm
    say("SpMM k=32 row-major, ordered: device")
    got = _device_spmm(ctx, m, a, b, c0, H.ALG_ORDERED)
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_cfg2_spmm_k32_plain_every_entry(ctx, cfg2, cfg2_spmm):
    """The plain form on the full matrix (B is 1 GiB: four 256 MiB column
    passes), every C(i,j) within 4*(len_i+1)*eps*(0.5*(|A||B|)_ij + |C0_ij|)
    of the serial result, with (|A||B|)_ij bounded above by |A|*m, m_c the
    largest |B(c,j)| of row c of B."""
    m, a, _ = cfg2
    b, c0, ref = cfg2_spmm
    say("SpMM k=32 row-major, plain: device")
    got = _device_spmm(ctx, m, a, b, c0, H.ALG_AUTO)
    say("SpMM k=32 row-major, plain: bound")
    rowmax = np.abs(b).max(axis=1)
    mag = O.abs_spmv_rows(m.offsets, m.indices, m.values, rowmax)
    lens = np.diff(m.offsets).astype(np.float64)
    bound = 4 * (lens[:, None] + 1) * EPS64 * (0.5 * mag[:, None] + np.abs(c0))
    err = np.abs(got - ref)
    worst = float(np.max(err / np.maximum(bound, 1e-300)))
    print(f"SpMM plain: max error / bound {worst:.3e}", flush=True)
    assert np.all(err <= bound), worst


def test_cfg2_spmm_k32_panel_every_entry(ctx, cfg2, cfg2_spmm):
    """The column-panel form on the full matrix (stream-staged analysis,
    16384-column panels, rows of at least 4096 entries listed: 12.9M
    segments): every C(i,j) within the plain bound, repeatable bit for bit."""
    m, a, _ = cfg2
    b, c0, ref = cfg2_spmm
    say("SpMM k=32 row-major, panel form: analysis")
    H.csr_spmv_analyze(ctx, a, H.panel_options(panel_columns=16384, long_row_min=4096,
                                                staging=H.STAGING_STREAM))
    say("SpMM k=32 row-major, panel form: device, twice")
    got = _device_spmm(ctx, m, a, b, c0, H.ALG_PANEL)
    again = _device_spmm(ctx, m, a, b, c0, H.ALG_PANEL)
    a.analysis.free()
    a.analysis = None
    H.plan_clear(ctx)
    np.testing.assert_array_equal(again.view(np.uint64), got.view(np.uint64))
    del again
    say("SpMM k=32 row-major, panel form: bound")
    rowmax = np.abs(b).max(axis=1)
    mag = O.abs_spmv_rows(m.offsets, m.indices, m.values, rowmax)
    lens = np.diff(m.offsets).astype(np.float64)
    bound = 4 * (lens[:, None] + 1) * EPS64 * (0.5 * mag[:, None] + np.abs(c0))
    err = np.abs(got - ref)
    worst = float(np.max(err / np.maximum(bound, 1e-300)))
    print(f"SpMM panel: max error / bound {worst:.3e}", flush=True)
    assert np.all(err <= bound), worst


def _row_of_entry(m):
    return np.repeat(np.arange(m.rows, dtype=np.int32), np.diff(m.offsets))


def _per_column_sums(keys, weights, offsets_c, sorted_vals):
    """Per-column sums of `weights` grouped by CSR column `keys` (bincount,
    exact: every sum stays below 2^53) and of `sorted_vals` over the CSC
    column ranges (reduceat in integers)."""
    cols = offsets_c.shape[0] - 1
    by_csr = np.bincount(keys, weights=weights, minlength=cols)
    starts = offsets_c[:-1]
    nonempty = np.diff(offsets_c) > 0
    by_csc = np.zeros(cols, dtype=np.uint64)
    by_csc[nonempty] = np.add.reduceat(sorted_vals.astype(np.uint64), starts[nonempty])
    return by_csr, by_csc


def test_cfg2_csr_to_csc_and_back(ctx, cfg2):
    m, a, _ = cfg2
    say("CSR->CSC on the device")
    csc = H.convert_compressed(ctx, a)
    off_c, idx_c, val_c = H.download_compressed(ctx, csc)
    counts = np.bincount(m.indices, minlength=m.cols)
    say("CSC: per-column counts and canonical order")
    assert off_c[0] == 0 and off_c[-1] == m.nnz
    np.testing.assert_array_equal(np.diff(off_c), counts)
    col_start = np.zeros(m.nnz, dtype=bool)
    col_start[off_c[:-1][counts > 0]] = True
    step = np.diff(idx_c.astype(np.int64))
    assert np.all((step > 0) | col_start[1:]), "rows do not strictly increase within a column"
    assert idx_c.min() >= 0 and idx_c.max() < m.rows
    say("CSC: per-column checksums of row ids and value bits")
    rows_of = _row_of_entry(m)
    by_csr, by_csc = _per_column_sums(m.indices, rows_of, off_c, idx_c)
    np.testing.assert_array_equal(by_csr.astype(np.uint64), by_csc)
    del rows_of
    bits, bits_c = m.values.view(np.uint64), val_c.view(np.uint64)
    for shift in (0, 32):
        w = ((bits >> np.uint64(shift)) & np.uint64(0xFFFFFFFF)).astype(np.float64)
        by_csr, by_csc = _per_column_sums(m.indices, w, off_c,
                                          (bits_c >> np.uint64(shift)) & np.uint64(0xFFFFFFFF))
        np.testing.assert_array_equal(by_csr.astype(np.uint64), by_csc)
        del w
    say("CSC: sampled entries found with their exact bits")
    rng = np.random.default_rng(32)
    for p in rng.integers(0, m.nnz, 1 << 14):
        r = int(np.searchsorted(m.offsets, p, side="right") - 1)
        c = int(m.indices[p])
        lo, hi = int(off_c[c]), int(off_c[c + 1])
        q = lo + int(np.searchsorted(idx_c[lo:hi], r))
        assert q < hi and idx_c[q] == r and bits_c[q] == bits[p], (p, r, c)
    del off_c, idx_c, val_c, bits_c
    say("CSC->CSR on the device: the original bit for bit")
    back = H.convert_compressed(ctx, csc)
    del csc
    off_b, idx_b, val_b = H.download_compressed(ctx, back)
    del back
    np.testing.assert_array_equal(off_b, m.offsets)
    np.testing.assert_array_equal(idx_b.astype(np.int32), m.indices)
    np.testing.assert_array_equal(val_b.view(np.uint64), m.values.view(np.uint64))


def test_cfg2_coordinate_round_trip(ctx, cfg2):
    m, a, _ = cfg2
    say("CSR->COO on the device")
    coo = H.convert_to_coordinate(ctx, a)
    coords = ctx.download(coo.buffers[0], np.int64, 2 * m.nnz)  # interleaved (row, col)
    np.testing.assert_array_equal(coords[0::2], _row_of_entry(m))
    np.testing.assert_array_equal(coords[1::2], m.indices)
    del coords
    vals = ctx.download(coo.buffers[1], np.float64, m.nnz)
    np.testing.assert_array_equal(vals.view(np.uint64), m.values.view(np.uint64))
    del vals
    say("COO->CSR on the device: the original bit for bit")
    csr = H.convert_coordinate(ctx, coo, H.CSR, 4)
    del coo
    off_b, idx_b, val_b = H.download_compressed(ctx, csr)
    del csr
    np.testing.assert_array_equal(off_b, m.offsets)
    np.testing.assert_array_equal(idx_b, m.indices)
    np.testing.assert_array_equal(val_b.view(np.uint64), m.values.view(np.uint64))


def test_cfg2_shuffled_triplets_through_finalize(ctx, cfg2):
    m, _, _ = cfg2
    say("shuffled COO: seeded permutation of the entries")
    perm = np.random.default_rng(0x5EEDA5C1).permutation(m.nnz)
    coords = np.empty((m.nnz, 2), dtype=np.int64)
    coords[:, 0] = _row_of_entry(m)[perm]
    coords[:, 1] = m.indices[perm]
    vals = m.values[perm]
    del perm
    t = H.triplets(ctx, m.rows, m.cols, coords, vals, H.F64)
    del coords, vals
    say("shuffled COO: device finalize (stable sort, sum duplicates, keep zeros)")
    fin = H.coo_finalize(ctx, t, H.DUPLICATE_SUM, H.ZERO_KEEP)
    del t
    assert fin.nnz == m.nnz
    say("shuffled COO: to CSR, the original bit for bit")
    csr = H.convert_coordinate(ctx, fin, H.CSR, 4)
    del fin
    off_b, idx_b, val_b = H.download_compressed(ctx, csr)
    del csr
    np.testing.assert_array_equal(off_b, m.offsets)
    np.testing.assert_array_equal(idx_b, m.indices)
    np.testing.assert_array_equal(val_b.view(np.uint64), m.values.view(np.uint64))


@pytest.fixture(scope="module")
def batch10k():
    say("batched: generating 10,000 matrices")
    specs = G.batched_specs(10_000)
    p = G.pack([G.make_spec(s) for s in specs])
    n = int(p.rows.sum())
    say(f"batched: {n} rows, {int(p.nnz.sum())} entries")
    return p, n, G.vector(n, 1), G.vector(n, 2)


def _batch_reference(p, x, y0, alpha, beta, dt):
    n = y0.shape[0]
    ref = y0.astype(dt).copy()
    absax = np.empty(n)
    lens = np.empty(n, dtype=np.int64)
    vals = p.values.astype(dt)
    xd = x.astype(dt)
    vabs = vals.astype(np.float64)
    xabs = xd.astype(np.float64)
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        r, o, z, s = int(p.rows[i]), int(p.off_start[i]), int(p.nz_start[i]), int(p.vec_start[i])
        off = p.offsets[o:o + r + 1]
        idx, v = p.indices[z:z + int(p.nnz[i])], vals[z:z + int(p.nnz[i])]
        O.compat_spmv(off, idx, v, alpha, xd[s:s + r], beta, ref[s:s + r])
        O.abs_spmv(off, idx, vabs[z:z + int(p.nnz[i])], xabs[s:s + r], absax[s:s + r])
        lens[s:s + r] = np.diff(off)
    with ThreadPoolExecutor(O.threads()) as ex:
        list(ex.map(one, range(p.rows.shape[0])))
    return ref, absax, lens


@pytest.mark.parametrize("kind", [H.F64, H.F32])
def test_batched_10k_graph_replay_every_matrix(ctx, batch10k, kind):
    p, n, x, y0 = batch10k
    dt, eps = (np.float64, EPS64) if kind == H.F64 else (np.float32, EPS32)
    alpha, beta = 1.25, -0.5
    say(f"batched 10k {dt.__name__}: oracle")
    ref, absax, lens = _batch_reference(p, x, y0, alpha, beta, dt)
    entries, bufs = H.upload_packed(ctx, p, x, y0, kind, alpha, beta)
    y0d = y0.astype(dt)
    b = H.Batch(ctx, entries, keep=bufs)
    say(f"batched 10k {dt.__name__}: tune, capture")
    b.tune(reps=2)
    b.capture()
    for rep in range(2):
        ctx.copy(bufs[4].ptr, y0d.ctypes.data, y0d.nbytes)  # y0 again (beta != 0)
        ctx.synchronize()
        say(f"batched 10k {dt.__name__}: graph replay {rep}")
        b.replay()
        ctx.synchronize()
        y = ctx.download(bufs[4], dt, n)
        bound_check(y, ref, lens, absax, alpha, beta, y0d.astype(np.float64), eps,
                    f"batched replay {rep}")
    del b, entries, bufs


def test_precision_full_size_against_fp64(ctx):
    say("precision workload n=2^20: oracle")
    m = G.precision_matrix(1 << 20)
    x64 = G.vector(m.cols, 4242)
    y64 = O.compat_spmv_rows(m.offsets, m.indices, m.values, 1.0, x64, 0.0, np.zeros(m.rows))
    norm = float(np.max(O.abs_spmv_rows(m.offsets, m.indices, m.values, x64)))
    a64 = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    xb = ctx.upload(x64.astype(np.float32))
    yb = ctx.alloc(4 * m.rows)
    vmax = float(np.max(np.abs(m.values)))
    for kind, tol in ((H.F32, 4e-6), (H.BF16, 1e-2), (H.F16, 1.5e-3)):
        scale = 1.0
        while kind == H.F16 and vmax * scale > 32768.0:
            scale *= 0.5
        ak = H.convert_values(ctx, a64, kind, scale)
        for alg, acc in ((H.ALG_VECTOR[16], H.ACC_PLAIN), (H.ALG_VECTOR[16], H.ACC_TWOSUM),
                         (H.ALG_MERGE, H.ACC_PLAIN)):
            H.csr_spmv(ctx, 1.0 / scale, ak, H.vector(xb, m.cols, H.F32, writable=False), 0.0,
                       H.vector(yb, m.rows, H.F32), alg=alg, acc=acc)
            y = ctx.download(yb, np.float32, m.rows).astype(np.float64)
            err = float(np.max(np.abs(y - y64)) / norm)
            say(f"precision kind {kind} {H.ALG_NAMES[alg]} acc {acc}: rel. error {err:.3e}")
            assert np.isfinite(y).all() and err <= tol
        H.csr_spmv_analyze(ctx, ak, H.panel_options(staging=H.STAGING_STREAM))
        H.csr_spmv(ctx, 1.0 / scale, ak, H.vector(xb, m.cols, H.F32, writable=False), 0.0,
                   H.vector(yb, m.rows, H.F32), alg=H.ALG_PANEL)
        y = ctx.download(yb, np.float32, m.rows).astype(np.float64)
        err = float(np.max(np.abs(y - y64)) / norm)
        say(f"precision kind {kind} stream plan: rel. error {err:.3e}")
        assert np.isfinite(y).all() and err <= tol
        ak.analysis.free()
        ak.analysis = None
        H.plan_clear(ctx)
