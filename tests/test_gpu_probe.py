"""The gather-rate probe computes what it states: sum over p of
values[p] * x[min(indices[p], x_extent - 1)] (partials summed on the host,
checked to a summation-order bound), including indices past the extent,
which it clamps."""
import numpy as np
import pytest

import asc_sparse_hip as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,xn", [(1, 1), (1000, 7), (300_000, 1 << 15), (1 << 20, 1 << 22)])
def test_probe_sum(ctx, n, xn):
    rng = np.random.default_rng(n + xn)
    idx = rng.integers(0, xn + 5, n).astype(np.int32)  # a few past the end: clamped
    val = rng.uniform(-1, 1, n)
    x = rng.uniform(-1, 1, xn)
    bi, bv, bx = ctx.upload(idx), ctx.upload(val), ctx.upload(x)
    out = ctx.alloc(8 * H.GATHER_PROBE_BLOCKS)
    ctx.fill(out, 0)
    print(f"probe n {n} x {xn}", flush=True)
    H.gather_probe(ctx, bi, bv, n, bx, xn, out)
    got = ctx.download(out, np.float64, H.GATHER_PROBE_BLOCKS).sum()
    prods = val * x[np.minimum(idx, xn - 1)]
    assert abs(got - prods.sum()) <= 4 * n * 2.0 ** -53 * np.abs(prods).sum() + 1e-300
