"""Batched SpMV: one persistent launch over many uneven CSR matrices, eager
and hipGraph-replayed, against the oracle run matrix by matrix.

The plain batched kernel reassociates per row, so each y_i is held to the
plain-mode bound 4*len*eps*|alpha|*(|A||x|)_i + 4*eps*|beta*y0_i|.  Graph
replays are checked against the eager result after every replay, with eager
work and synchronizations in between.
"""
import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
import oracle_bind as O

pytestmark = pytest.mark.gpu


def build_batch(ctx, count, kind, seed=7, hi=4096, alpha=1.25, beta=-0.5):
    specs = G.batched_specs(count, seed=seed, hi=hi)
    mats = [G.make_spec(s) for s in specs]
    p = G.pack(mats)
    n = int(p.rows.sum())
    x = G.vector(n, 3)
    y0 = G.vector(n, 4)
    entries, bufs = H.upload_packed(ctx, p, x, y0, kind, alpha, beta)
    return mats, p, x, y0, entries, bufs


def reference(mats, p, x, y0, alpha, beta, dt):
    out = np.empty_like(y0, dtype=np.float64)
    bound = np.empty_like(out)
    eps = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    for i, m in enumerate(mats):
        s, r = int(p.vec_start[i]), m.rows
        xs = x[s:s + r].astype(dt)
        ys = y0[s:s + r].astype(dt).copy()
        O.compat_spmv(m.offsets, m.indices, m.values.astype(dt), alpha, xs, beta, ys)
        out[s:s + r] = ys
        absax = O.abs_spmv(m.offsets, m.indices, m.values.astype(dt).astype(np.float64), xs.astype(np.float64))
        lens = np.diff(m.offsets)
        bound[s:s + r] = 4 * np.maximum(lens, 1) * eps * abs(alpha) * absax + 4 * eps * np.abs(beta * y0[s:s + r])
    return out, bound


@pytest.mark.parametrize("kind", [H.F64, H.F32])
def test_batched_matches_oracle(ctx, kind):
    dt = np.float64 if kind == H.F64 else np.float32
    mats, p, x, y0, entries, bufs = build_batch(ctx, 300, kind)
    ref, bound = reference(mats, p, x, y0, 1.25, -0.5, dt)
    b = H.Batch(ctx, entries, keep=bufs)
    b.run()
    y = ctx.download(bufs[4], dt, y0.shape[0]).astype(np.float64)
    assert np.all(np.abs(y - ref) <= bound)


def test_graph_replays_match_eager(ctx):
    mats, p, x, y0, entries, bufs = build_batch(ctx, 200, H.F64, seed=11, alpha=0.5, beta=0.0)
    b = H.Batch(ctx, entries, keep=bufs)
    b.run()
    eager = ctx.download(bufs[4], np.float64, y0.shape[0])
    b.capture()
    for rep in range(4):
        ctx.fill(bufs[4], 0xFF)  # poison y (NaN pattern) between replays
        ctx.synchronize()
        b.replay()
        ctx.synchronize()
        got = ctx.download(bufs[4], np.float64, y0.shape[0])
        np.testing.assert_array_equal(got.view(np.uint64), eager.view(np.uint64), err_msg=f"replay {rep}")
        b.run()  # eager work between replays
        ctx.synchronize()


def test_class_lanes_do_not_change_results_beyond_bound(ctx):
    mats, p, x, y0, entries, bufs = build_batch(ctx, 120, H.F64, seed=5)
    ref, bound = reference(mats, p, x, y0, 1.25, -0.5, np.float64)
    b = H.Batch(ctx, entries, keep=bufs)
    y0_dev = ctx.upload(y0)
    # 0 selects CSR-stream items
    for lanes in ([1] * 8, [64] * 8, [4, 8, 2, 16, 1, 32, 64, 8], [0] * 8, [0, 1, 0, 8, 0, 32, 0, 64]):
        ctx.copy(bufs[4].ptr, y0_dev.ptr, y0.nbytes)  # restore y0 before every run
        ctx.synchronize()
        b.build(lanes)
        b.run()
        y = ctx.download(bufs[4], np.float64, y0.shape[0])
        assert np.all(np.abs(y - ref) <= bound), lanes


def test_tuner_returns_valid_lanes(ctx):
    mats, p, x, y0, entries, bufs = build_batch(ctx, 90, H.F64, seed=9)
    b = H.Batch(ctx, entries, keep=bufs)
    lanes, ms = b.tune(reps=2)
    assert all(g in (0, 1, 2, 4, 8, 16, 32, 64) for g in lanes)
    present = [c for c in range(8) if any(t >= 0 for t in ms[c])]
    assert present, "no class was timed"
    for c in present:
        slot = 7 if lanes[c] == 0 else lanes[c].bit_length() - 1  # timings: lanes 1..64, stream
        assert ms[c][slot] == min(t for t in ms[c] if t >= 0)


def test_stream_items_that_overflow_the_window(ctx):
    """Rows far longer than their matrix's average: a CSR-stream item holding
    one of them exceeds the 2048-entry LDS window and takes the lane path."""
    rng = np.random.default_rng(21)
    mats = []
    for rows, long_rows in ((1000, (0, 517)), (3000, (2999,)), (700, ())):
        lens = np.ones(rows, dtype=np.int64)
        for r in long_rows:
            lens[r] = 2500 + 300 * r % 400
        cols = rows + 3000
        offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        idx = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.int32)
        mats.append(G.Csr(rows, cols, offsets, idx, rng.uniform(-1, 1, offsets[-1])))
    p = G.pack(mats)
    n = int(p.rows.sum())
    # square vectors per matrix are what upload_packed lays out: give each
    # matrix an x as long as its columns by padding x with the extra columns
    xs = [G.vector(m.cols, 30 + i) for i, m in enumerate(mats)]
    y0 = G.vector(n, 40)
    ref, bound = [], []
    for i, m in enumerate(mats):
        ys = y0[p.vec_start[i]:p.vec_start[i] + m.rows].copy()
        O.compat_spmv(m.offsets, m.indices, m.values, 0.75, xs[i], 0.25, ys)
        absax = O.abs_spmv(m.offsets, m.indices, m.values, xs[i])
        ref.append(ys)
        bound.append(4 * np.maximum(np.diff(m.offsets), 1) * 2.0 ** -53 * 0.75 * absax
                     + 4 * 2.0 ** -53 * np.abs(0.25 * y0[p.vec_start[i]:p.vec_start[i] + m.rows]))
    bo, bi, bv = ctx.upload(p.offsets), ctx.upload(p.indices), ctx.upload(p.values)
    bx = [ctx.upload(v) for v in xs]
    by = ctx.upload(y0)
    entries = []
    for i, m in enumerate(mats):
        a = H.SparseDesc(bo.ptr + 8 * int(p.off_start[i]), bi.ptr + 4 * int(p.nz_start[i]),
                         bv.ptr + 8 * int(p.nz_start[i]), m.rows, m.cols, m.nnz, H.CSR, H.F64, 4, 8,
                         H.DEVICE, 1)
        entries.append(H.BatchEntry(a, H.vector(bx[i], m.cols, H.F64, writable=False),
                                    H.vector(by.ptr + 8 * int(p.vec_start[i]), m.rows, H.F64), 0.75, 0.25))
    b = H.Batch(ctx, entries, keep=[bo, bi, bv, by] + bx, lanes=[0] * 8)
    print("stream items with window overflow", flush=True)
    b.run()
    y = ctx.download(by, np.float64, n)
    for i, m in enumerate(mats):
        got = y[p.vec_start[i]:p.vec_start[i] + m.rows]
        assert np.all(np.abs(got - ref[i]) <= bound[i]), i


def test_rejects_overlapping_outputs(ctx):
    mats, p, x, y0, entries, bufs = build_batch(ctx, 4, H.F64, seed=2)
    entries[1].y.data = entries[0].y.data  # two matrices writing the same y
    with pytest.raises(H.AscHipError) as e:
        H.Batch(ctx, entries, keep=bufs)
    assert e.value.code == H.INVALID_ARGUMENT
