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


# ---- the two-pass transpose (inner extents 2^9..2^22) -----------------------
# Column extents of 2^9..2^22 take the two-pass radix transpose: rows derived
# in the tile, (high digit, row) packed in 32 bits with the row's top bits
# recovered from row-segment split points when they do not fit, column
# offsets read from the second pass's scanned counts.  These cases aim at
# those parts: dropped row bits (2 and 4 of them), tiles spanning millions of
# empty rows, one column spread over many tiles, narrow values.

def sparse_rows_csr(rows, cols, nonempty, per_row_max, seed, dense_col=None, vdtype=np.float64):
    """CSR with `nonempty` random rows of 1..per_row_max distinct random
    columns (plus `dense_col` in every row when given); values random bits."""
    rng = np.random.default_rng(seed)
    which = np.sort(rng.choice(rows, size=nonempty, replace=False)) if nonempty < rows else np.arange(rows)
    lens = rng.integers(1, per_row_max + 1, size=which.shape[0])
    rid = np.repeat(which, lens)
    cid = rng.integers(0, cols, size=rid.shape[0])
    if dense_col is not None:
        rid = np.concatenate([rid, which])
        cid = np.concatenate([cid, np.full(which.shape[0], dense_col)])
    order = np.lexsort((cid, rid))
    rid, cid = rid[order], cid[order]
    keep = np.ones(rid.shape[0], dtype=bool)
    keep[1:] = (rid[1:] != rid[:-1]) | (cid[1:] != cid[:-1])
    rid, cid = rid[keep], cid[keep]
    off = np.zeros(rows + 1, dtype=np.int64)
    np.add.at(off, rid + 1, 1)
    off = np.cumsum(off)
    nbytes = np.dtype(vdtype).itemsize
    vals = rng.integers(0, 1 << (8 * nbytes - 1), size=rid.shape[0], dtype=np.int64).astype(
        {8: np.uint64, 4: np.uint32, 2: np.uint16}[nbytes]).view(vdtype)
    return off, cid.astype(np.int64), vals


def check_transpose_both_ways(ctx, rows, cols, off, idx, vals, kind, ib):
    ro, ri, rv = O.transpose(0, rows, cols, off, idx, vals)
    csr = H.clone_compressed(ctx, H.CSR, rows, cols, off, idx, vals, kind, ib)
    csc = H.convert_compressed(ctx, csr)
    check_compressed(ctx, csc, ro, ri, rv)
    back = H.convert_compressed(ctx, csc)
    check_compressed(ctx, back, off, idx, vals)


@pytest.mark.parametrize("rows,cols,nonempty,per_row", [
    ((1 << 22) + 17, (1 << 21) + 3, 2_000_000, 4),  # 23 row bits + 11-bit digit: 2 bits dropped
    ((1 << 24) + 5, (1 << 21) + 9, 2_500_000, 3),   # 25 row bits: 4 bits dropped (16 segments)
])
@pytest.mark.parametrize("ib", [4, 8])
def test_two_pass_row_segments(ctx, rows, cols, nonempty, per_row, ib):
    off, idx, vals = sparse_rows_csr(rows, cols, nonempty, per_row, seed=rows % 1009)
    print(f"two-pass segments rows {rows} cols {cols} nnz {vals.shape[0]} ib {ib}", flush=True)
    check_transpose_both_ways(ctx, rows, cols, off, idx, vals, H.F64, ib)


@pytest.mark.parametrize("kind", [H.F64, H.F32, H.BF16])
def test_two_pass_dense_column(ctx, kind):
    # Column 777777 is in every row: its low-digit run spans ten second-pass tiles.
    rows, cols = 300_000, 1 << 20
    vd = {H.F64: np.float64, H.F32: np.float32, H.BF16: np.uint16}[kind]
    off, idx, vals = sparse_rows_csr(rows, cols, rows, 4, seed=5, dense_col=777_777, vdtype=vd)
    print(f"two-pass dense column kind {kind} nnz {vals.shape[0]}", flush=True)
    for ib in (4, 8):
        check_transpose_both_ways(ctx, rows, cols, off, idx, vals, kind, ib)


def test_two_pass_empty_row_gaps(ctx):
    # 20000 non-empty rows among 5M: tiles span millions of empty rows.
    rows, cols = 5_000_000, 4096
    off, idx, vals = sparse_rows_csr(rows, cols, 20_000, 5, seed=9)
    print(f"two-pass empty-row gaps nnz {vals.shape[0]}", flush=True)
    check_transpose_both_ways(ctx, rows, cols, off, idx, vals, H.F64, 4)


@pytest.mark.parametrize("cols", [257, 300, 511, 512, 513, 1 << 22])
def test_two_pass_extent_edges(ctx, cols):
    # Smallest and largest inner extents of the two-pass form, and the
    # digit-width steps in between.
    rows = 120_000
    off, idx, vals = sparse_rows_csr(rows, cols, rows - 1000, 6, seed=cols)
    print(f"two-pass extent cols {cols} nnz {vals.shape[0]}", flush=True)
    check_transpose_both_ways(ctx, rows, cols, off, idx, vals, H.F64, 8)
