"""GPU checks of the variant machinery: features, MLP kernels (against the
FP64 numpy oracle), training, the exhaustive autotuner, selection with its
merge-path fallback and its column-panel plan outputs, and -- on the
committed MI355X dataset -- the learned selector's regret against exhaustive
autotuning over every variant and against no network.
"""
import glob
import os

import numpy as np
import pytest

import asc_sparse_hip as H
from asc_sparse_hip import gen as G
from asc_sparse_hip import selector as S
import mlp_oracle as M

pytestmark = pytest.mark.gpu

DATA_DIR = os.path.join(os.path.dirname(S.__file__), "models")


@pytest.mark.parametrize("i", [0, 1, 2, 3, 4, 5, 10])
def test_features_match_oracle(ctx, i):
    sp = S.spec("heldout", 1000 + i)
    sp["rows"] = min(sp["rows"], 20000)
    if sp["family"] in (G.STENCIL2D, G.STENCIL3D):
        k = sp["grid"] = 20 if sp["family"] == G.STENCIL2D else 7
        sp["rows"] = k * k if sp["family"] == G.STENCIL2D else k ** 3
    if "zipf_max" in sp:
        sp["zipf_max"] = min(sp["zipf_max"], sp["rows"])
    m = S.make(sp)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    got = H.features(ctx, a)
    ref = M.features(m.offsets, m.indices, m.rows, m.cols)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-6)


def test_mlp_forward_and_gradient_match_oracle(ctx):
    rng = np.random.default_rng(1)
    k, b = 8, 37
    p = S.init_params(k, seed=5)
    x = rng.standard_normal((b, 16)).astype(np.float32)
    t = rng.standard_normal((b, k)).astype(np.float32)
    net = S.DeviceMLP(ctx, p, b, k)
    xb, tb = ctx.upload(x), ctx.upload(t)
    y = net.forward(xb, b)
    np.testing.assert_allclose(y, M.forward(p, x.astype(np.float64), k)[2], rtol=1e-4, atol=1e-5)
    net.gradient(xb, tb)
    g = ctx.download(net.g, np.float32, net.n)
    loss, gref = M.loss_and_grad(p, x, t, k)
    np.testing.assert_allclose(net.last_loss(), loss, rtol=1e-5)
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-6)


def test_training_reduces_loss_and_is_deterministic(ctx):
    rng = np.random.default_rng(2)
    feats = rng.standard_normal((200, 16)).astype(np.float32)
    times = np.exp(rng.standard_normal((200, 8)) * 0.3 + feats[:, :8] * 0.5).astype(np.float32)
    m1 = S.train(ctx, feats, times, steps=600, log=lambda *a, **k: None)
    m2 = S.train(ctx, feats, times, steps=600, log=lambda *a, **k: None)
    curve = m1["train"]["loss_curve"]
    assert curve[-1][1] < 0.5 * curve[0][1]
    assert m1["params"] == m2["params"]


def test_autotune_and_plan_cache(ctx):
    m = G.make(G.ZIPF, 30000, 30000, seed=4, values=G.VALUES_NORMAL, zipf_max=2000)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    xb, yb = ctx.upload(G.vector(m.cols, 1)), ctx.alloc(8 * m.rows)
    xd, yd = H.vector(xb, m.cols, H.F64, writable=False), H.vector(yb, m.rows, H.F64)
    best, times = H.autotune(ctx, a, xd, yd, reps=2)
    assert best in H.TUNABLE_ALGS and all(t > 0 for t in times.values())
    assert times[best] == min(times.values())
    assert H.plan_get(ctx, a) == best
    H.plan_clear(ctx)
    assert H.plan_get(ctx, a) == H.ALG_AUTO


def test_select_api_sets_plan_and_falls_back(ctx):
    rng = np.random.default_rng(3)
    feats = rng.standard_normal((64, 16)).astype(np.float32)
    times = np.exp(rng.standard_normal((64, 8))).astype(np.float32)
    model = S.train(ctx, feats, times, steps=50, log=lambda *a, **k: None)
    m = G.uniform16(4096)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    for clip, margin, expect_merge in ((1e-3, 0.0, True), (1e30, 1e30, True), (1e30, -1e30, False)):
        model["clip"], model["margin"] = clip, margin
        pick, scores = S.select(ctx, model, a)
        if expect_merge:
            assert pick.algorithm == H.ALG_MERGE and pick.fell_back == 1
        else:
            assert pick.algorithm == H.ALG_MERGE + int(np.argmin(scores)) and pick.fell_back == 0
        assert pick.staging == 0 and pick.direct_algorithm == pick.algorithm
        assert H.plan_get(ctx, a) == pick.algorithm


def test_select_with_plan_outputs(ctx):
    """An 11-output selector whose targets make a plan the best variant picks
    ALG_PANEL with that plan's staging, and leaves the best direct kernel in
    the plan cache; the caller's analysis with that staging then registers
    the plan."""
    rng = np.random.default_rng(4)
    m = G.uniform16(4096)
    a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
    # Training inputs around this matrix's own features, so that its
    # normalized features lie inside the training range.
    f0 = H.features(ctx, a)
    feats = (f0[None, :] * (1 + 0.05 * rng.standard_normal((96, 16)))).astype(np.float32)
    cost = np.exp(rng.uniform(0.5, 1.5, (96, 11)))
    cost[:, 2] = 1.2   # output 2 (vec2): the best direct kernel
    cost[:, 10] = 1.0  # the stream plan: the best variant
    model = S.train(ctx, feats, cost, steps=400, log=lambda *a, **k: None)
    assert model["k"] == 11
    model["clip"], model["margin"] = 1e30, -1e30
    pick, scores = S.select(ctx, model, a)
    assert len(scores) == 11 and int(np.argmin(scores)) == 10
    assert pick.algorithm == H.ALG_PANEL and pick.staging == H.STAGING_STREAM
    assert pick.direct_algorithm == H.ALG_MERGE + int(np.argmin(scores[:8])) == H.ALG_VECTOR[2]
    assert H.plan_get(ctx, a) == pick.direct_algorithm
    H.csr_spmv_analyze(ctx, a, H.panel_options(staging=pick.staging))
    assert H.plan_get(ctx, a) == H.ALG_PANEL
    H.plan_clear(ctx)


@pytest.mark.skipif(not glob.glob(os.path.join(DATA_DIR, "heldout*.npz")) or
                    not os.path.exists(S.DEFAULT_MODEL), reason="no committed selector dataset/model")
def test_learned_selector_regret_on_heldout(ctx):
    """On the 500 held-out matrices, against the exhaustive autotuner over all
    eleven variants (direct kernels and column-panel plans with amortized
    analysis): the selector's geometric-mean regret is at most 10%, below
    merge-path alone, the best single variant and the exhaustive autotuner
    over the direct kernels only."""
    d = S.load_data(sorted(glob.glob(os.path.join(DATA_DIR, "heldout*.npz"))))
    model = S.load_model()
    r = S.evaluate(ctx, model, d["features"], S.costs(d))
    print(r)
    assert r["outputs"] == model["k"] == 11
    assert r["regret_selector"] <= 0.10
    assert r["regret_selector"] < r["regret_merge_only"]
    assert r["regret_selector"] <= r["regret_best_single"]
    assert r["regret_selector"] < r["regret_direct_autotuner"]
