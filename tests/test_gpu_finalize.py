"""GPU parity of the device COO finalize (CoordinateBuilder::Finalize,
coordinate.h:1300-1497) through the C ABI, bit-exact against the oracle
(itself pinned to the reference's own builder).

Checks: the golden vectors of coordinate_test.cc:96-142,250-272,329-351;
random triplets with many duplicates (sums depend on insertion order) in FP64
and FP32 under every policy pair, at shapes that need one, two and three
radix passes per key; out-of-range coordinates; the reject path leaving the
outputs untouched; empty input; a 2^22-triplet case.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import asc_sparse_hip as H
import oracle_bind as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))
NP = {H.F64: np.float64, H.F32: np.float32}


def fetch(ctx, out):
    d = out.desc
    n = d.nnz
    coords = ctx.download(out.buffers[0], np.int64, 2 * n) if n else np.zeros(0, np.int64)
    vals = ctx.download(out.buffers[1], NP[d.element_kind], n) if n else np.zeros(0, NP[d.element_kind])
    return coords, vals


def same_values(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    np.testing.assert_array_equal(np.nan_to_num(a).view(np.uint8), np.nan_to_num(b).view(np.uint8))


def run_both(ctx, rows, cols, coords, vals, kind, dup, zero):
    t = H.triplets(ctx, rows, cols, coords, vals, kind)
    rc, oc, ov = O.finalize(coords, vals.astype(NP[kind]), dup, zero)
    if rc != 0:
        with pytest.raises(H.AscHipError) as e:
            H.coo_finalize(ctx, t, dup, zero)
        assert e.value.code == H.INVALID_ARGUMENT
        return None
    out = H.coo_finalize(ctx, t, dup, zero)
    assert out.desc.trusted == 1 and out.desc.nnz == ov.shape[0]
    gc, gv = fetch(ctx, out)
    np.testing.assert_array_equal(gc, oc)
    same_values(gv, ov)
    return out


@pytest.mark.parametrize("kind", [H.F64, H.F32])
def test_golden_vectors(ctx, kind):
    g = GOLD["finalize"]
    coords = np.array(g["coords"], np.int64)
    vals = np.array(g["values"], NP[kind])
    out = H.coo_finalize(ctx, H.triplets(ctx, 3, 4, coords, vals, kind), H.DUPLICATE_SUM, H.ZERO_KEEP)
    gc, gv = fetch(ctx, out)
    np.testing.assert_array_equal(gc, g["sum_keep_coords"])
    np.testing.assert_array_equal(gv, np.array(g["sum_keep_values"], NP[kind]))
    with pytest.raises(H.AscHipError) as e:
        H.coo_finalize(ctx, H.triplets(ctx, 3, 4, coords, vals, kind), H.DUPLICATE_REJECT, H.ZERO_KEEP)
    assert e.value.code == H.INVALID_ARGUMENT
    out = H.coo_finalize(ctx, H.triplets(ctx, 3, 4, coords, vals, kind), H.DUPLICATE_SUM, H.ZERO_DROP)
    assert out.desc.nnz == g["sum_drop_nnz"]
    nd = g["nan_drop"]
    nan_vals = np.array([float(v) for v in nd["values"]], NP[kind])
    out = H.coo_finalize(ctx, H.triplets(ctx, 1, 3, np.array(nd["coords"]), nan_vals, kind),
                         H.DUPLICATE_SUM, H.ZERO_DROP)
    assert out.desc.nnz == nd["nnz"]
    gc, gv = fetch(ctx, out)
    assert np.isnan(gv[0])


@pytest.mark.parametrize("kind", [H.F64, H.F32])
@pytest.mark.parametrize("shape", [(9, 7), (300, 1000), (70000, 300), (5, 1 << 17), (1 << 17, 1 << 17)])
@pytest.mark.parametrize("dup,zero", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_random_against_oracle(ctx, kind, shape, dup, zero):
    rows, cols = shape
    seed = rows * 31 + cols + 7 * dup + zero
    rng = np.random.default_rng(seed)
    n = 20000
    if dup == 1:
        # a small coordinate pool: many repeats, summed in insertion order
        pool = rng.integers(0, rows * cols, max(1, n // 8))
        flat = rng.choice(pool, n)
    else:
        flat = np.unique(rng.integers(0, rows * cols, n))
        rng.shuffle(flat)
        if seed % 3 == 0 and flat.size > 1:
            flat[-1] = flat[0]  # one duplicate: the reject path
    r, c = flat // cols, flat % cols
    vals = np.where(rng.random(flat.size) < 0.4, rng.integers(-2, 3, flat.size),
                    rng.standard_normal(flat.size)).astype(NP[kind])
    vals[::97] = np.nan
    coords = np.stack([r, c], 1).reshape(-1)
    print(f"finalize shape {shape} kind {kind} dup {dup} zero {zero} n {flat.size}", flush=True)
    run_both(ctx, rows, cols, coords, vals, kind, dup, zero)


def test_out_of_range_and_untouched_outputs(ctx):
    lib, h = ctx.lib, ctx.handle
    coords = np.array([0, 1, 2, 5, 1, 1], np.int64)  # (2, 5) is outside 3x4
    vals = np.array([1.0, 2.0, 3.0])
    t = H.triplets(ctx, 3, 4, coords, vals, H.F64)
    need = C.c_size_t()
    H.check(lib.asc_sparse_hip_coo_finalize_workspace_size(h, C.byref(t.desc), C.byref(need)))
    ws = ctx.alloc(max(need.value, 256))
    oc = ctx.upload(np.full(6, -7, np.int64))
    ov = ctx.upload(np.full(3, 9.5))
    out = H.SparseDesc()
    st = lib.asc_sparse_hip_coo_finalize(h, C.byref(t.desc), H.DUPLICATE_SUM, H.ZERO_KEEP, oc.ptr,
                                         ov.ptr, C.byref(out), ws.ptr, ws.nbytes)
    assert st.code == H.INDEX
    # duplicates under reject: outputs untouched as well
    t2 = H.triplets(ctx, 3, 4, np.array([0, 1, 2, 3, 0, 1], np.int64), vals, H.F64)
    st = lib.asc_sparse_hip_coo_finalize(h, C.byref(t2.desc), H.DUPLICATE_REJECT, H.ZERO_KEEP,
                                         oc.ptr, ov.ptr, C.byref(out), ws.ptr, ws.nbytes)
    assert st.code == H.INVALID_ARGUMENT
    np.testing.assert_array_equal(ctx.download(oc, np.int64, 6), np.full(6, -7))
    np.testing.assert_array_equal(ctx.download(ov, np.float64, 3), np.full(3, 9.5))


def test_empty(ctx):
    out = H.coo_finalize(ctx, H.triplets(ctx, 4, 4, np.zeros(0, np.int64), np.zeros(0), H.F64),
                         H.DUPLICATE_SUM, H.ZERO_DROP)
    assert out.desc.nnz == 0 and out.desc.trusted == 1


def test_large_then_convert_to_csr(ctx):
    """2^22 triplets over a 2^20 x 2^20 shape (three radix passes per key),
    then the device COO -> CSR of the result against the oracle's."""
    rng = np.random.default_rng(5)
    n, dim = 1 << 22, 1 << 20
    r = rng.integers(0, dim, n)
    c = rng.integers(0, dim, n)
    r[1::5], c[1::5] = r[::5][: r[1::5].size], c[::5][: c[1::5].size]  # guaranteed repeats
    vals = rng.standard_normal(n)
    coords = np.stack([r, c], 1).reshape(-1)
    out = run_both(ctx, dim, dim, coords, vals, H.F64, H.DUPLICATE_SUM, H.ZERO_DROP)
    csr = H.convert_coordinate(ctx, out, H.CSR, 8)
    gc, gv = fetch(ctx, out)
    rows = gc[0::2]
    off = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=dim))])
    np.testing.assert_array_equal(ctx.download(csr.buffers[0], np.int64, dim + 1), off)
    np.testing.assert_array_equal(ctx.download(csr.buffers[1], np.int64, out.desc.nnz), gc[1::2])
