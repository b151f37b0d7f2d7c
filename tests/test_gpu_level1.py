"""GPU parity of the sparse Level-1 operations (blas.h:1330-1582) through the
C ABI: the golden vectors of standard_blas_test.cc:69-174; random canonical
indexed vectors against the oracle (itself pinned to the reference), bit-exact
for axpy, gather, gather-zero, scatter and the ordered dot, and within
nnz*eps*sum|v*x| for the plain (tree) dot; strided and reversed dense vectors;
alpha == 0 with NaN values; an empty indexed vector.
"""
import json
import os

import numpy as np
import pytest

import asc_sparse_hip as H
import oracle_bind as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
KIND = {np.float64: H.F64, np.float32: H.F32}
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}


def dense_view(ctx, storage, extent, stride, kind, writable=True):
    """Device copy of `storage`; logical element i at storage[i*stride]
    (negative strides address from the far end)."""
    buf = ctx.upload(storage)
    esz = storage.itemsize
    first = buf.ptr + (0 if stride > 0 else (storage.shape[0] - 1) * esz)
    return buf, H.vector(first, extent, kind, stride, writable=writable)


def logical(storage, extent, stride):
    return storage[: (extent - 1) * stride + 1: stride] if stride > 0 else storage[::stride][:extent]


def run(ctx, op, idx, v, storage, extent, stride, dtype, alpha=1.0, alg=H.ALG_AUTO):
    kind = KIND[dtype]
    sp = H.clone_indexed(ctx, idx, v.astype(dtype), extent, kind)
    buf, dv = dense_view(ctx, storage.astype(dtype), extent, stride, kind)
    res = ctx.upload(np.full(1, 7.0, dtype))
    rv = H.vector(res, 1, kind)
    H.level1(ctx, op, sp, dv, alpha, rv if op == H.LEVEL1_DOT else None, alg)
    dot = ctx.download(res, dtype, 1)[0]
    vals = ctx.download(sp.buffers[1], dtype, idx.shape[0]) if idx.shape[0] else np.zeros(0, dtype)
    st = ctx.download(buf, dtype, storage.shape[0])
    return dot, vals, st


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_golden_vectors(ctx, dtype):
    g = GOLD["level1"]
    idx, v, x = np.array(g["indices"]), np.array(g["values"], dtype), np.array(g["dense"], dtype)
    for alg in (H.ALG_AUTO, H.ALG_ORDERED):
        assert run(ctx, H.LEVEL1_DOT, idx, v, x, 3, 1, dtype, alg=alg)[0] == g["dot"]
    np.testing.assert_array_equal(run(ctx, H.LEVEL1_AXPY, idx, v, x, 3, 1, dtype, g["axpy_alpha"])[2], g["axpy"])
    np.testing.assert_array_equal(run(ctx, H.LEVEL1_GATHER, idx, np.zeros(2), x, 3, 1, dtype)[1], g["gather"])
    _, gv, gx = run(ctx, H.LEVEL1_GATHER_ZERO, idx, np.zeros(2), x, 3, 1, dtype)
    np.testing.assert_array_equal(gv, g["gather_zero_values"])
    np.testing.assert_array_equal(gx, g["gather_zero_dense"])
    np.testing.assert_array_equal(run(ctx, H.LEVEL1_SCATTER, idx, v, np.zeros(3), 3, 1, dtype)[2],
                                  g["scatter_into_zeros"])
    # alpha == 0: untouched even with NaN values (blas.h:1388-1390)
    out = run(ctx, H.LEVEL1_AXPY, idx, np.full(2, np.nan), x, 3, 1, dtype, 0.0)[2]
    np.testing.assert_array_equal(out, x)


OPS = {H.LEVEL1_AXPY: "axpy", H.LEVEL1_GATHER: "gather", H.LEVEL1_GATHER_ZERO: "gather_zero",
       H.LEVEL1_SCATTER: "scatter"}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("stride", [1, 3, -1, -2])
@pytest.mark.parametrize("size", [(0, 10), (1, 1), (700, 5000), (200000, 1 << 20)])
def test_random_against_oracle(ctx, dtype, stride, size):
    nnz, extent = size
    rng = np.random.default_rng(nnz * 7 + abs(stride))
    idx = np.sort(rng.choice(extent, nnz, replace=False)).astype(np.int64)
    v = rng.standard_normal(nnz).astype(dtype)
    storage = rng.standard_normal((extent - 1) * abs(stride) + 1).astype(dtype)
    x = np.ascontiguousarray(logical(storage, extent, stride))
    print(f"level1 {dtype.__name__} stride {stride} nnz {nnz} extent {extent}", flush=True)
    for op, name in OPS.items():
        _, gv, gst = run(ctx, op, idx, v, storage, extent, stride, dtype, alpha=-1.75)
        _, ov, ox = O.level1(name, idx, v if op != H.LEVEL1_GATHER else v, x, 1, -1.75)
        if op in (H.LEVEL1_GATHER, H.LEVEL1_GATHER_ZERO):
            np.testing.assert_array_equal(gv.view(np.uint8), ov.view(np.uint8))
        np.testing.assert_array_equal(np.ascontiguousarray(logical(gst, extent, stride)).view(np.uint8),
                                      ox.view(np.uint8))
        holes = np.ones(storage.shape[0], bool)
        holes[np.arange(extent) * abs(stride) if stride > 0 else
              storage.shape[0] - 1 - np.arange(extent) * abs(stride)] = False
        np.testing.assert_array_equal(gst[holes], storage[holes])  # strided holes untouched
    ref = O.level1("dot", idx, v, x)[0]
    got = run(ctx, H.LEVEL1_DOT, idx, v, storage, extent, stride, dtype, alg=H.ALG_ORDERED)[0]
    assert np.asarray([got], dtype).tobytes() == np.asarray([ref], dtype).tobytes()
    plain = run(ctx, H.LEVEL1_DOT, idx, v, storage, extent, stride, dtype)[0]
    bound = max(nnz, 1) * EPS[dtype] * float(np.sum(np.abs(v.astype(np.float64) * x[idx].astype(np.float64))))
    assert abs(float(plain) - float(ref)) <= bound + 1e-30
