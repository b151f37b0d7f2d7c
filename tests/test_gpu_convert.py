"""GPU conversions through the C ABI: canonical COO -> CSR/CSC and CSR <-> CSC
transposes, bit-exact on offsets, indices and values against the oracle (and
the reference's golden vectors).  Full-size checks use size-independent
properties: transpose twice is the identity, offsets equal column counts.
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


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def check_compressed(ctx, dev, off, idx, val):
    o, i, v = H.download_compressed(ctx, dev)
    np.testing.assert_array_equal(o, off)
    np.testing.assert_array_equal(i.astype(np.int64), np.asarray(idx, dtype=np.int64))
    np.testing.assert_array_equal(bits(v), bits(val))


def test_golden_all_directions(ctx):
    g = GOLD["conversion"]
    rows, cols = g["shape"]
    coo = H.clone_coordinate(ctx, rows, cols, np.array(g["coo"]), np.array(g["coo_values"], float), H.F64)
    for ib in (4, 8):
        csr = H.convert_coordinate(ctx, coo, H.CSR, ib)
        check_compressed(ctx, csr, g["csr_offsets"], g["csr_indices"], np.array(g["csr_values"], float))
        csc = H.convert_coordinate(ctx, coo, H.CSC, ib)
        check_compressed(ctx, csc, g["csc_offsets"], g["csc_indices"], np.array(g["csc_values"], float))
        t = H.convert_compressed(ctx, csr)
        check_compressed(ctx, t, g["csc_offsets"], g["csc_indices"], np.array(g["csc_values"], float))
        t2 = H.convert_compressed(ctx, csc)
        check_compressed(ctx, t2, g["csr_offsets"], g["csr_indices"], np.array(g["csr_values"], float))


@pytest.mark.parametrize("shape", GOLD["conversion"]["empty_shapes"])
def test_empty_shapes(ctx, shape):
    rows, cols = shape
    coo = H.clone_coordinate(ctx, rows, cols, np.zeros(0, np.int64), np.zeros(0), H.F64)
    for fmt in (H.CSR, H.CSC):
        out = H.convert_coordinate(ctx, coo, fmt, 8)
        outer = rows if fmt == H.CSR else cols
        check_compressed(ctx, out, [0] * (outer + 1), [], np.zeros(0))
    csr = H.clone_compressed(ctx, H.CSR, rows, cols, np.zeros(rows + 1, np.int64), np.zeros(0, np.int64),
                             np.zeros(0), H.F64)
    check_compressed(ctx, H.convert_compressed(ctx, csr), [0] * (cols + 1), [], np.zeros(0))


def test_explicit_zero_and_signed_zero_survive(ctx):
    off, idx = np.array([0, 1, 3]), np.array([0, 0, 1])
    vals = np.array([0.0, -0.0, 3.0])
    csr = H.clone_compressed(ctx, H.CSR, 2, 2, off, idx, vals, H.F64)
    t = H.convert_compressed(ctx, csr)
    ro, ri, rv = O.transpose(0, 2, 2, off, idx, vals)
    check_compressed(ctx, t, ro, ri, rv)


@pytest.mark.parametrize("name,rows,cols", [("zipf", 1 << 16, 1 << 16), ("wide", 3000, 300000),
                                             ("tall", 200000, 700), ("tiny_cols", 5000, 1)])
@pytest.mark.parametrize("kind", [H.F64, H.F32])
def test_random_transpose_bit_exact(ctx, name, rows, cols, kind):
    zmax = min(cols, 10000)
    m = G.make(G.ZIPF, rows, cols, seed=len(name) * 1000 + rows % 997, values=G.VALUES_NORMAL, zipf_max=zmax)
    val = m.values if kind == H.F64 else m.values.astype(np.float32)
    ro, ri, rv = O.transpose(0, rows, cols, m.offsets, m.indices, val)
    for ib in (4, 8):
        print(f"transpose {name} kind {kind} ib {ib} nnz {m.nnz}", flush=True)
        csr = H.clone_compressed(ctx, H.CSR, rows, cols, m.offsets, m.indices, val, kind, ib)
        csc = H.convert_compressed(ctx, csr)
        check_compressed(ctx, csc, ro, ri, rv)
        back = H.convert_compressed(ctx, csc)
        check_compressed(ctx, back, m.offsets, m.indices, val)


@pytest.mark.parametrize("kind", [H.F64, H.F32])
def test_coo_to_csr_and_csc_random(ctx, kind):
    m = G.make(G.ZIPF, 40000, 50000, seed=77, values=G.VALUES_NORMAL, zipf_max=5000)
    val = m.values if kind == H.F64 else m.values.astype(np.float32)
    coords, cval = O.compressed_to_coo(0, m.rows, m.offsets, m.indices, val)
    coo = H.clone_coordinate(ctx, m.rows, m.cols, coords, cval, kind)
    for fmt in (0, 1):
        ro, ri, rv = O.coo_to_compressed(fmt, m.rows, m.cols, coords, cval)
        for ib in (4, 8):
            out = H.convert_coordinate(ctx, coo, H.CSR if fmt == 0 else H.CSC, ib)
            check_compressed(ctx, out, ro, ri, rv)


def test_transpose_many_digits(ctx):
    # Column extent needing three 8-bit radix rounds and a sparse pattern.
    rows, cols = 3000, 5_000_000
    m = G.make(G.UNIFORM_FIXED, rows, cols, seed=3, values=G.VALUES_UNIFORM, per_row=300)
    ro, ri, rv = O.transpose(0, rows, cols, m.offsets, m.indices, m.values)
    csr = H.clone_compressed(ctx, H.CSR, rows, cols, m.offsets, m.indices, m.values, H.F64, 4)
    check_compressed(ctx, H.convert_compressed(ctx, csr), ro, ri, rv)


@pytest.mark.parametrize("fmt", [H.CSR, H.CSC])
def test_compressed_to_coordinate(ctx, fmt):
    g = GOLD["conversion"]
    rows, cols = g["shape"]
    off = g["csr_offsets"] if fmt == H.CSR else g["csc_offsets"]
    idx = g["csr_indices"] if fmt == H.CSR else g["csc_indices"]
    val = np.array(g["csr_values"] if fmt == H.CSR else g["csc_values"], float)
    a = H.clone_compressed(ctx, fmt, rows, cols, np.array(off), np.array(idx), val, H.F64)
    coo = H.convert_to_coordinate(ctx, a)
    coords = ctx.download(H._PtrView(coo.desc.outer_offsets), np.int64, 2 * coo.nnz)
    vals = ctx.download(H._PtrView(coo.desc.values), np.float64, coo.nnz)
    assert coords.tolist() == g["coo"]
    np.testing.assert_array_equal(vals, g["coo_values"])


def test_compressed_to_coordinate_random_round_trip(ctx):
    m = G.make(G.ZIPF, 30000, 20000, seed=12, values=G.VALUES_NORMAL, zipf_max=3000)
    rc, rv = O.compressed_to_coo(0, m.rows, m.offsets, m.indices, m.values)
    for fmt in (H.CSR, H.CSC):
        if fmt == H.CSR:
            a = H.clone_compressed(ctx, H.CSR, m.rows, m.cols, m.offsets, m.indices, m.values, H.F64, 4)
        else:
            co, ci, cv = O.transpose(0, m.rows, m.cols, m.offsets, m.indices, m.values)
            a = H.clone_compressed(ctx, H.CSC, m.rows, m.cols, co, ci, cv, H.F64, 4)
        coo = H.convert_to_coordinate(ctx, a)
        coords = ctx.download(H._PtrView(coo.desc.outer_offsets), np.int64, 2 * coo.nnz)
        vals = ctx.download(H._PtrView(coo.desc.values), np.float64, coo.nnz)
        np.testing.assert_array_equal(coords, rc)
        np.testing.assert_array_equal(vals.view(np.uint64), rv.view(np.uint64))
        back = H.convert_coordinate(ctx, coo, H.CSR, 4)
        check_compressed(ctx, back, m.offsets, m.indices, m.values)
