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
    for lanes in ([1] * 8, [64] * 8, [4, 8, 2, 16, 1, 32, 64, 8]):
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
    assert all(g in (1, 2, 4, 8, 16, 32, 64) for g in lanes)
    present = [c for c in range(8) if any(t >= 0 for t in ms[c])]
    assert present, "no class was timed"
    for c in present:
        assert ms[c][lanes[c].bit_length() - 1] == min(t for t in ms[c] if t >= 0)


def test_rejects_overlapping_outputs(ctx):
    mats, p, x, y0, entries, bufs = build_batch(ctx, 4, H.F64, seed=2)
    entries[1].y.data = entries[0].y.data  # two matrices writing the same y
    with pytest.raises(H.AscHipError) as e:
        H.Batch(ctx, entries, keep=bufs)
    assert e.value.code == H.INVALID_ARGUMENT
