"""CPU checks of the selector's oracle and host logic.

The numpy MLP oracle's analytic gradient is pinned by central finite
differences; the feature oracle by hand-computed values on tiny matrices;
the fallback rule and the regret measure by constructed cases.
"""
import numpy as np
import pytest

import mlp_oracle as M
from asc_sparse_hip import selector as S


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    k = 8
    p = S.init_params(k, seed=3).astype(np.float64)
    x = rng.standard_normal((13, 16))
    t = rng.standard_normal((13, k))
    loss, g = M.loss_and_grad(p, x, t, k)
    idx = rng.choice(p.size, 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (M.loss_and_grad(p + e, x, t, k)[0] - M.loss_and_grad(p - e, x, t, k)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_param_count_and_layout():
    assert S.init_params(8).size == 64 * 16 + 64 + 64 * 64 + 64 + 8 * 64 + 8


def test_feature_oracle_on_tiny_matrices():
    # 3x4: rows of length 2, 2, 1 (golden CSR of the reference's tests)
    f = M.features([0, 2, 4, 5], [1, 3, 0, 2, 1], 3, 4)
    assert f[0] == 2.0 and f[1] == np.log2(5) and f[2] == np.log2(6)
    assert np.isclose(f[3], np.log2(1 + 5 / 3))
    assert f[6] == 0.0 and np.isclose(f[7], 1.0)  # all lengths in [1, 4)
    # column gaps within rows: row0 3-1=2, row1 2-0=2
    assert np.isclose(f[15], np.log2(3.0))
    # empty rows count
    f = M.features([0, 0, 1, 1], [0], 3, 3)
    assert np.isclose(f[6], 2 / 3)


def test_choose_falls_back_when_unsure():
    pred = np.array([[0.30, 0.00, 0.5], [0.01, 0.00, 0.5], [0.5, 0.0, 0.2]])
    norm = np.zeros((3, 16))
    norm[2, 5] = 50.0  # out of the training range
    pick = S.choose(pred, norm, margin=0.05, clip=6.0)
    assert pick.tolist() == [1, 0, 0]


def test_regret_is_zero_for_oracle_choice_and_positive_otherwise():
    times = np.array([[1.0, 2.0], [3.0, 1.5]])
    assert S.regret(times, np.array([0, 1])) == pytest.approx(0.0)
    assert S.regret(times, np.array([1, 0])) == pytest.approx(np.sqrt(2.0 * 2.0) - 1.0)


def test_specs_are_deterministic_and_cover_six_families():
    names = {S.spec("train", i)["name"].replace("stencil2d", "stencil").replace("stencil3d", "stencil")
             for i in range(12)}
    assert names >= {"uniform", "zipf", "banded", "blockdiag8", "stencil", "bipartite"}
    assert S.spec("train", 7) == S.spec("train", 7)
    assert S.spec("train", 7) != S.spec("heldout", 7)
    for i in range(60):
        r = S.spec("train", i)["rows"]
        assert S.ROWS_LO // 2 <= r <= S.ROWS_HI * 1.1


def test_costs_amortize_plan_analysis_and_targets_cap_refusals():
    d = {"times": np.array([[2.0, 4.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0]], np.float32),
         "plan_spmv": np.array([[0.5, np.inf, 0.25]], np.float32),
         "plan_analysis": np.array([[100.0, np.inf, 50.0]], np.float32)}
    c = S.costs(d)
    assert c.shape == (1, S.K) and S.K == 11
    # plan cost = SpMV + analysis / 100
    assert c[0, 8] == pytest.approx(0.5 + 1.0) and c[0, 10] == pytest.approx(0.25 + 0.5)
    assert np.isinf(c[0, 9])
    t = S.targets(c)
    assert t[0, 10] == pytest.approx(0.0)           # the best variant
    assert t[0, 2] == pytest.approx(np.log(1.0 / 0.75))
    assert t[0, 9] == S.TARGET_CAP                  # refused plan: capped
    assert S.OUTPUT_NAMES[8:] == ["panel_l2", "panel_lds", "panel_stream"]


def test_regret_caps_refused_picks():
    cost = np.array([[1.0, np.inf]])
    assert S.regret(cost, np.array([1])) == pytest.approx(1e6 - 1.0)
