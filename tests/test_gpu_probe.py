"""The measurement probes compute what they state.  Gather: sum over p of
values[p] * x[min(indices[p], x_extent - 1)] (partials summed on the host,
checked to a summation-order bound), including indices past the extent,
which it clamps.  Stream read: the wrapping sum of every 32-bit word.
Row gather: the sum of every element of table row min(indices[p],
table_rows - 1) over p.  Stream copy: the bytes, unchanged."""
import numpy as np
import pytest

from asc_sparse_hip import probe as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,xn", [(1, 1), (1000, 7), (300_000, 1 << 15), (1 << 20, 1 << 22)])
def test_probe_sum(ctx, n, xn):
    rng = np.random.default_rng(n + xn)
    idx = rng.integers(0, xn + 5, n).astype(np.int32)  # a few past the end: clamped
    val = rng.uniform(-1, 1, n)
    x = rng.uniform(-1, 1, xn)
    bi, bv, bx = ctx.upload(idx), ctx.upload(val), ctx.upload(x)
    out = ctx.alloc(8 * P.GATHER_PROBE_BLOCKS)
    ctx.fill(out, 0)
    print(f"probe n {n} x {xn}", flush=True)
    P.gather_probe(ctx, bi, bv, n, bx, xn, out)
    got = ctx.download(out, np.float64, P.GATHER_PROBE_BLOCKS).sum()
    prods = val * x[np.minimum(idx, xn - 1)]
    assert abs(got - prods.sum()) <= 4 * n * 2.0 ** -53 * np.abs(prods).sum() + 1e-300


@pytest.mark.parametrize("words", [1, 1000, 1 << 20, (1 << 22) + 7])
def test_stream_probes(ctx, words):
    rng = np.random.default_rng(words)
    data = rng.integers(0, 2 ** 32, 4 * words, dtype=np.uint64).astype(np.uint32)
    src, dst = ctx.upload(data), ctx.alloc(data.nbytes)
    out = ctx.alloc(4 * P.STREAM_PROBE_BLOCKS)
    ctx.fill(out, 0)
    print(f"stream probes, {words} words", flush=True)
    P.stream_read(ctx, src, data.nbytes, out)
    got = int(ctx.download(out, np.uint32, P.STREAM_PROBE_BLOCKS).astype(np.uint64).sum()) % 2 ** 32
    assert got == int(data.astype(np.uint64).sum()) % 2 ** 32
    P.stream_copy(ctx, src, dst, data.nbytes)
    np.testing.assert_array_equal(ctx.download(dst, np.uint32, data.size), data)


@pytest.mark.parametrize("n,rows", [(1, 1), (7, 3), (1000, 17), (300_001, 1 << 12)])
def test_row_gather_probe_sum(ctx, n, rows):
    rng = np.random.default_rng(n + rows)
    idx = rng.integers(0, rows + 3, n).astype(np.int32)  # a few past the end: clamped
    table = rng.uniform(-1, 1, (rows, 32))
    bi, bt = ctx.upload(idx), ctx.upload(table)
    out = ctx.alloc(8 * P.ROW_GATHER_PROBE_BLOCKS)
    ctx.fill(out, 0)
    print(f"row gather probe n {n} rows {rows}", flush=True)
    P.row_gather_probe(ctx, bi, n, bt, rows, out)
    got = ctx.download(out, np.float64, P.ROW_GATHER_PROBE_BLOCKS).sum()
    gathered = table[np.minimum(idx, rows - 1)]
    assert abs(got - gathered.sum()) <= 4 * 32 * n * 2.0 ** -53 * np.abs(gathered).sum() + 1e-300
