"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against (a) the reference's own known-answer
vectors (tests/golden/reference_golden.json), (b) what mathematics fixes
independently of the oracle's formula (exact integer/dyadic arithmetic,
dense brute force, Python's stable sorted(), transpose involution), and (c)
the reference compiled from its own sources (oracle/_ref) when present.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_bind as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_golden.json")))


def arr(v, dtype=np.float64):
    return np.array([float(t) if t != "nan" else np.nan for t in v], dtype=dtype)


# ---- compat SpMV: golden vectors -------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(GOLD["compat_spmv"])))
def test_compat_spmv_golden(case, dtype):
    g = GOLD["compat_spmv"][case]
    vals = arr(g["values"], dtype)
    x = arr(g.get("x_storage", g.get("x")), dtype)
    y = arr(g["y_storage"], dtype)
    O.compat_spmv(g["offsets"], g["indices"], vals, g["alpha"], x, g["beta"], y,
                  incx=g["incx"], incy=g["incy"])
    if "expected_storage" in g:
        np.testing.assert_array_equal(y, arr(g["expected_storage"], dtype))
    else:
        np.testing.assert_array_equal(y[:: g["incy"]][: g["rows"]], arr(g["expected_rows"], dtype))


def test_compat_spmv_left_to_right_order():
    g = GOLD["compat_order_f32"]
    y = arr(g["y0"], np.float32)
    O.compat_spmv(g["offsets"], g["indices"], arr(g["values"], np.float32), g["alpha"],
                  arr(g["x"], np.float32), g["beta"], y)
    assert y[0] == np.float32(0.0)


def test_compat_spmv_beta_zero_does_not_read_y():
    # NaN in y must not leak when beta == 0 (reference_blas.cc:35-37).
    y = np.full(3, np.nan)
    O.compat_spmv([0, 1, 1, 2], [0, 1], np.array([1.0, 2.0]), 1.0, np.array([3.0, 4.0]), 0.0, y)
    np.testing.assert_array_equal(y, [3.0, 0.0, 8.0])


def test_compat_spmv_empty_row_negative_alpha_gives_negative_zero():
    y = np.array([5.0])
    O.compat_spmv([0, 0], [], np.array([], dtype=np.float64), -2.0, np.array([1.0]), 0.0, y)
    assert y[0] == 0.0 and np.signbit(y[0])


def test_compat_spmv_exact_dyadic_bruteforce():
    # Dyadic values whose every partial sum is exactly representable: the
    # result is fixed by exact rational arithmetic, independent of the oracle.
    rng = np.random.default_rng(7)
    rows, cols = 200, 150
    dense = np.where(rng.random((rows, cols)) < 0.1, rng.integers(-64, 64, (rows, cols)) / 8.0, 0.0)
    off = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int64)
    idx = np.nonzero(dense)[1].astype(np.int64)
    val = dense[dense != 0]
    x = rng.integers(-16, 16, cols) / 4.0
    y0 = rng.integers(-16, 16, rows) / 2.0
    alpha, beta = 0.75, -1.5
    y = y0.copy()
    O.compat_spmv(off, idx, val, alpha, x, beta, y)
    for r in range(rows):
        exact = sum(Fraction(dense[r, c]) * Fraction(x[c]) for c in range(cols)) * Fraction(alpha) \
            + Fraction(beta) * Fraction(y0[r])
        assert Fraction(y[r]) == exact


def test_compat_spmv_unfused_rounding():
    # p = fl(-1 + fl(v1*x1)): the product (1+2^-30)^2 = 1 + 2^-29 + 2^-60
    # rounds to 1 + 2^-29 first, so the sum is exactly 2^-29.  A fused
    # multiply-add would keep the 2^-60 term.
    v = np.array([-1.0, 1.0 + 2.0 ** -30])
    x = np.array([1.0, 1.0 + 2.0 ** -30])
    y = np.zeros(1)
    O.compat_spmv([0, 2], [0, 1], v, 1.0, x, 0.0, y)
    assert y[0] == 2.0 ** -29


# ---- |A||x|: the quantity every plain-mode tolerance scales ---------------------------

def _sparsity(rng, rows, cols, density):
    mask = rng.random((rows, cols)) < density
    off = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int64)
    idx = np.nonzero(mask)[1].astype(np.int64)
    return mask, off, idx


def test_abs_spmv_exact_integers():
    # Integer operands of mixed sign: every product and partial sum is an
    # integer below 2^53, so the result is exact and must equal the integer
    # sum of |a_ij| * |x_j|.
    rng = np.random.default_rng(21)
    rows, cols = 300, 170  # rectangular: a transposed operand fails
    mask, off, idx = _sparsity(rng, rows, cols, 0.2)
    dense = np.where(mask, rng.integers(-1000, 1000, (rows, cols)), 0)
    val = dense[mask].astype(np.float64)
    x = rng.integers(-1000, 1000, cols).astype(np.float64)
    out = O.abs_spmv(off, idx, val, x)
    expect = np.abs(dense).astype(np.int64) @ np.abs(x).astype(np.int64)
    np.testing.assert_array_equal(out, expect.astype(np.float64))


def test_abs_spmv_correctly_rounded_within_one_ulp():
    # Random doubles over many binades: the exact rational sum of |v||x|,
    # rounded once (float(Fraction) rounds to nearest), must agree with the
    # oracle to one unit in the last place.
    rng = np.random.default_rng(22)
    rows, cols = 120, 90
    mask, off, idx = _sparsity(rng, rows, cols, 0.3)
    val = rng.standard_normal(idx.size) * np.exp2(rng.integers(-30, 30, idx.size))
    x = rng.uniform(-1, 1, cols) * np.exp2(rng.integers(-20, 20, cols))
    out = O.abs_spmv(off, idx, val, x)
    for r in range(rows):
        exact = sum((abs(Fraction(float(val[p]))) * abs(Fraction(float(x[idx[p]])))
                     for p in range(off[r], off[r + 1])), Fraction(0))
        ref = float(exact)
        assert out[r] >= 0.0
        assert abs(out[r] - ref) <= np.spacing(ref), (r, out[r], ref)


def test_abs_spmv_is_not_the_plain_product():
    # Cancelling signs: A x is exactly zero while |A||x| is not, so a dropped
    # absolute value (on v or on x) gives a different answer.
    off = np.array([0, 2, 4], dtype=np.int64)
    idx = np.array([0, 1, 0, 1], dtype=np.int64)
    val = np.array([3.0, 3.0, -2.0, 2.0])
    x = np.array([1.5, -1.5])
    np.testing.assert_array_equal(O.abs_spmv(off, idx, val, x), [9.0, 6.0])
    y = np.zeros(2)
    O.compat_spmv(off, idx, val, 1.0, x, 0.0, y)
    np.testing.assert_array_equal(y, [0.0, -6.0])


def test_narrow_index_variants_equal_wide():
    # The *_i32 entry points run the same loops over int32 indices; chunked
    # calls over row ranges (the threaded helpers) give the same bits.
    rng = np.random.default_rng(23)
    rows, cols = 3000, 2500
    mask, off, idx = _sparsity(rng, rows, cols, 0.01)
    val = rng.standard_normal(idx.size)
    x = rng.uniform(-1, 1, cols)
    y64, y32, yt = (G := rng.uniform(-1, 1, rows)).copy(), G.copy(), G.copy()
    O.compat_spmv(off, idx, val, -1.25, x, 0.5, y64)
    O.compat_spmv(off, idx.astype(np.int32), val, -1.25, x, 0.5, y32)
    O.compat_spmv_rows(off, idx.astype(np.int32), val, -1.25, x, 0.5, yt)
    np.testing.assert_array_equal(y64.view(np.uint64), y32.view(np.uint64))
    np.testing.assert_array_equal(y64.view(np.uint64), yt.view(np.uint64))
    f64 = O.compat_spmv(off, idx, val.astype(np.float32), np.float32(2.0), x.astype(np.float32),
                        0.0, np.zeros(rows, np.float32))
    f32 = O.compat_spmv(off, idx.astype(np.int32), val.astype(np.float32), np.float32(2.0),
                        x.astype(np.float32), 0.0, np.zeros(rows, np.float32))
    np.testing.assert_array_equal(f64.view(np.uint32), f32.view(np.uint32))
    a64 = O.abs_spmv(off, idx, val, x)
    np.testing.assert_array_equal(a64, O.abs_spmv(off, idx.astype(np.int32), val, x))
    np.testing.assert_array_equal(a64, O.abs_spmv_rows(off, idx.astype(np.int32), val, x))
    k = 5
    b = rng.uniform(-1, 1, (cols, k))
    c0 = rng.uniform(-1, 1, (rows, k))
    cw, cn = c0.copy(), c0.copy()
    O.std_spmm(0, 0, rows, cols, off, idx, val, 0.5, b, 1, k, k, cw, 1, k)
    O.std_spmm(0, 0, rows, cols, off, idx.astype(np.int32), val, 0.5, b, 1, k, k, cn, 1, k)
    np.testing.assert_array_equal(cw.view(np.uint64), cn.view(np.uint64))


# ---- standard SpMV / SpMM ----------------------------------------------------------

def test_standard_golden():
    g = GOLD["standard"]
    rows, cols = g["shape"]
    vals = arr(g["values"])
    x = arr(g["x"])
    for fmt, off, idx in ((0, g["csr_offsets"], g["csr_indices"]), (1, g["csc_offsets"], g["csc_indices"])):
        y = np.zeros(2)
        O.std_spmv(fmt, 0, rows, cols, off, idx, vals, 1.0, x, y)
        np.testing.assert_array_equal(y, arr(g["spmv_n"]))
        y = np.zeros(2)
        O.std_spmv(fmt, 2, rows, cols, off, idx, vals, 1.0, x, y)
        np.testing.assert_array_equal(y, arr(g["spmv_t"]))
    # alpha == 0: NaN matrix is never read, y unchanged.
    y = arr(g["nan_alpha_zero_y0"])
    O.std_spmv(0, 0, rows, cols, g["csr_offsets"], g["csr_indices"], np.full(3, np.nan), 0.0, x, y)
    np.testing.assert_array_equal(y, arr(g["nan_alpha_zero_y0"]))
    c = np.zeros(4)
    O.std_spmm(0, 0, rows, cols, g["csr_offsets"], g["csr_indices"], vals, 1.0,
               arr(g["spmm_row_major_b"]), 1, 2, 2, c, 1, 2)
    np.testing.assert_array_equal(c, arr(g["spmm_row_major_c"]))
    c = np.zeros(4)
    O.std_spmm(1, 2, rows, cols, g["csc_offsets"], g["csc_indices"], vals, 1.0,
               arr(g["spmm_csc_h_col_major_b"]), 0, 2, 2, c, 0, 2)
    np.testing.assert_array_equal(c, arr(g["spmm_csc_h_col_major_c"]))
    c = arr(g["spmm_alpha_zero_c0"])
    O.std_spmm(0, 0, rows, cols, g["csr_offsets"], g["csr_indices"], np.full(3, np.nan), 0.0,
               arr(g["spmm_row_major_b"]), 1, 2, 2, c, 1, 2)
    np.testing.assert_array_equal(c, arr(g["spmm_alpha_zero_c0"]))


def _random_csr(rng, rows, cols, density, ints=True):
    dense = np.where(rng.random((rows, cols)) < density,
                     rng.integers(-9, 10, (rows, cols)).astype(np.float64) if ints else rng.standard_normal((rows, cols)),
                     0.0)
    mask = rng.random((rows, cols)) < density
    dense = np.where(mask, dense, 0.0)
    nzmask = mask
    off = np.concatenate([[0], np.cumsum(nzmask.sum(1))]).astype(np.int64)
    idx = np.nonzero(nzmask)[1].astype(np.int64)
    val = dense[nzmask]
    return dense, off, idx, val


@pytest.mark.parametrize("op", [0, 1, 2])
def test_standard_spmv_dense_bruteforce(op):
    rng = np.random.default_rng(11 + op)
    rows, cols = 37, 53
    dense, off, idx, val = _random_csr(rng, rows, cols, 0.2)
    x = rng.integers(-5, 6, cols if op == 0 else rows).astype(np.float64)
    y0 = rng.integers(-5, 6, rows if op == 0 else cols).astype(np.float64)
    y = y0.copy()
    O.std_spmv(0, op, rows, cols, off, idx, val, 3.0, x, y)
    a = dense if op == 0 else dense.T
    np.testing.assert_array_equal(y, y0 + 3.0 * (a @ x))  # small integers: exact
    # CSC form of the same matrix gives the same result.
    co, ci, cv = O.transpose(0, rows, cols, off, idx, val)
    y2 = y0.copy()
    O.std_spmv(1, op, rows, cols, co, ci, cv, 3.0, x, y2)
    np.testing.assert_array_equal(y2, y)


@pytest.mark.parametrize("layout", [0, 1])
def test_standard_spmm_dense_bruteforce(layout):
    rng = np.random.default_rng(5)
    rows, cols, k = 29, 41, 7
    dense, off, idx, val = _random_csr(rng, rows, cols, 0.25)
    b = rng.integers(-4, 5, (cols, k)).astype(np.float64)
    c0 = rng.integers(-4, 5, (rows, k)).astype(np.float64)
    if layout == 1:
        bb, cc, ldb, ldc = b.reshape(-1).copy(), c0.reshape(-1).copy(), k, k
    else:
        bb, cc, ldb, ldc = b.T.reshape(-1).copy(), c0.T.reshape(-1).copy(), cols, rows
    O.std_spmm(0, 0, rows, cols, off, idx, val, -2.0, bb, layout, ldb, k, cc, layout, ldc)
    got = cc.reshape(rows, k) if layout == 1 else cc.reshape(k, rows).T
    np.testing.assert_array_equal(got, c0 - 2.0 * (dense @ b))


def test_std_spmv_per_entry_rounding_order():
    # y += (alpha*v)*x entry by entry from the old y (blas.h:1653-1665):
    # starting from y0 = 1e16 the +1 terms are each absorbed.
    y = np.array([1e16])
    O.std_spmv(0, 0, 1, 3, [0, 3], [0, 1, 2], np.array([1.0, 1.0, 1.0]), 1.0, np.ones(3), y)
    assert y[0] == 1e16
    # compat sums first from +0, then scales: the +1 terms survive as 3.
    y = np.array([1e16])
    O.compat_spmv([0, 3], [0, 1, 2], np.array([1.0, 1.0, 1.0]), 1.0, np.ones(3), 1.0, y)
    assert y[0] == 1e16 + 4.0  # 1e16 + 3 is a tie; round-to-even gives +4


# ---- conversions -----------------------------------------------------------------

def test_conversion_golden_all_directions():
    g = GOLD["conversion"]
    rows, cols = g["shape"]
    coo = np.array(g["coo"], dtype=np.int64)
    cv = arr(g["coo_values"])
    for faithful in (False, True):
        o, i, v = O.coo_to_compressed(0, rows, cols, coo, cv, faithful)
        assert o.tolist() == g["csr_offsets"] and i.tolist() == g["csr_indices"]
        np.testing.assert_array_equal(v, arr(g["csr_values"]))
        o, i, v = O.coo_to_compressed(1, rows, cols, coo, cv, faithful)
        assert o.tolist() == g["csc_offsets"] and i.tolist() == g["csc_indices"]
        np.testing.assert_array_equal(v, arr(g["csc_values"]))
        o, i, v = O.transpose(0, rows, cols, g["csr_offsets"], g["csr_indices"], arr(g["csr_values"]), faithful)
        assert o.tolist() == g["csc_offsets"] and i.tolist() == g["csc_indices"]
        np.testing.assert_array_equal(v, arr(g["csc_values"]))
        o, i, v = O.transpose(1, rows, cols, g["csc_offsets"], g["csc_indices"], arr(g["csc_values"]), faithful)
        assert o.tolist() == g["csr_offsets"] and i.tolist() == g["csr_indices"]
        np.testing.assert_array_equal(v, arr(g["csr_values"]))
    for fmt, off, idx, val in ((0, g["csr_offsets"], g["csr_indices"], g["csr_values"]),
                               (1, g["csc_offsets"], g["csc_indices"], g["csc_values"])):
        c, v = O.compressed_to_coo(fmt, len(off) - 1, off, idx, arr(val))
        assert c.tolist() == g["coo"]
        np.testing.assert_array_equal(v, cv)


@pytest.mark.parametrize("shape", GOLD["conversion"]["empty_shapes"])
def test_conversion_empty_shapes(shape):
    rows, cols = shape
    e = np.zeros(0)
    for fmt in (0, 1):
        o, i, v = O.coo_to_compressed(fmt, rows, cols, np.zeros(0, np.int64), e)
        assert o.tolist() == [0] * ((rows if fmt == 0 else cols) + 1) and i.size == 0
    o, i, v = O.transpose(0, rows, cols, np.zeros(rows + 1, np.int64), np.zeros(0, np.int64), e)
    assert o.tolist() == [0] * (cols + 1)


def test_conversion_keeps_explicit_zero_bits():
    g = GOLD["conversion"]["explicit_zero"]
    vals = np.array([-0.0, 3.0])  # a signed zero must survive bit-for-bit
    o, i, v = O.transpose(0, 2, 2, g["offsets"], g["indices"], vals)
    assert v.view(np.uint64).tolist() == vals.view(np.uint64).tolist()


def _random_pattern(rng, rows, cols, nnz):
    flat = rng.choice(rows * cols, size=nnz, replace=False)
    flat.sort()
    r, c = flat // cols, flat % cols
    return r.astype(np.int64), c.astype(np.int64)


@pytest.mark.parametrize("seed", range(4))
def test_fast_equals_faithful_and_sorted_bruteforce(seed):
    rng = np.random.default_rng(100 + seed)
    rows, cols = int(rng.integers(1, 60)), int(rng.integers(1, 60))
    nnz = int(rng.integers(0, rows * cols // 2 + 1))
    r, c = _random_pattern(rng, rows, cols, nnz)
    vals = rng.standard_normal(nnz)
    coo = np.stack([r, c], 1).reshape(-1)
    for fmt in (0, 1):
        fast = O.coo_to_compressed(fmt, rows, cols, coo, vals)
        faith = O.coo_to_compressed(fmt, rows, cols, coo, vals, True)
        for a, b in zip(fast, faith):
            np.testing.assert_array_equal(a, b)
        # Independent: Python's stable sort by the new outer coordinate.
        key = (lambda t: (t[0], t[1])) if fmt == 0 else (lambda t: (t[1], t[0]))
        trip = sorted(zip(r.tolist(), c.tolist(), vals.tolist()), key=key)
        inner = [t[1] if fmt == 0 else t[0] for t in trip]
        assert fast[1].tolist() == inner
        assert fast[2].tolist() == [t[2] for t in trip]
        outer = rows if fmt == 0 else cols
        counts = np.bincount(r if fmt == 0 else c, minlength=outer)
        assert fast[0].tolist() == [0] + np.cumsum(counts).tolist()
    csr = O.coo_to_compressed(0, rows, cols, coo, vals)
    t1 = O.transpose(0, rows, cols, *csr)
    t1f = O.transpose(0, rows, cols, *csr, faithful=True)
    for a, b in zip(t1, t1f):
        np.testing.assert_array_equal(a, b)
    back = O.transpose(1, rows, cols, *t1)  # involution
    for a, b in zip(back, csr):
        np.testing.assert_array_equal(a, b)


# ---- finalize --------------------------------------------------------------------

def test_finalize_golden():
    g = GOLD["finalize"]
    rc, c, v = O.finalize(g["coords"], g["values"], 1, 0)
    assert rc == 0 and c.tolist() == g["sum_keep_coords"] and v.tolist() == g["sum_keep_values"]
    rc, c, v = O.finalize(g["coords"], g["values"], 0, 0)
    assert rc == 1
    rc, c, v = O.finalize(g["coords"], g["values"], 1, 1)
    assert rc == 0 and v.shape[0] == g["sum_drop_nnz"]
    nd = g["nan_drop"]
    rc, c, v = O.finalize(nd["coords"], arr(nd["values"]), 0, 1)
    assert rc == 0 and v.shape[0] == nd["nnz"] and np.isnan(v[0])


def test_finalize_bruteforce():
    rng = np.random.default_rng(3)
    n = 300
    r = rng.integers(0, 7, n)
    c = rng.integers(0, 5, n)
    vals = rng.integers(-3, 4, n).astype(np.float64)
    rc, oc, ov = O.finalize(np.stack([r, c], 1).reshape(-1), vals, 1, 1)
    assert rc == 0
    groups = {}
    for i in range(n):  # insertion order within each coordinate
        groups.setdefault((int(r[i]), int(c[i])), []).append(vals[i])
    expect = []
    for key in sorted(groups):
        s = 0.0
        first = True
        for t in groups[key]:
            s = t if first else s + t
            first = False
        if s != 0.0:
            expect.append((key, s))
    assert [tuple(oc[2 * i: 2 * i + 2]) for i in range(len(ov))] == [e[0] for e in expect]
    assert ov.tolist() == [e[1] for e in expect]


def test_validate_rejects_noncanonical():
    assert O.validate(3, 4, [0, 2, 4, 5], [1, 3, 0, 2, 1], 5) == 0
    assert O.validate(3, 4, [0, 2, 4, 5], [3, 1, 0, 2, 1], 5) != 0  # unsorted
    assert O.validate(3, 4, [0, 2, 4, 5], [1, 3, 0, 2, 4], 5) != 0  # out of range
    assert O.validate(3, 4, [1, 2, 4, 5], [1, 3, 0, 2, 1], 5) != 0  # bad start


def test_level1_golden_vectors():
    g = GOLD["level1"]
    idx, v, x = np.array(g["indices"]), np.array(g["values"]), np.array(g["dense"])
    assert O.level1("dot", idx, v, x)[0] == g["dot"]
    np.testing.assert_array_equal(O.level1("axpy", idx, v, x, alpha=g["axpy_alpha"])[2], g["axpy"])
    np.testing.assert_array_equal(O.level1("gather", idx, np.zeros(2), x)[1], g["gather"])
    _, gv, gx = O.level1("gather_zero", idx, np.zeros(2), x)
    np.testing.assert_array_equal(gv, g["gather_zero_values"])
    np.testing.assert_array_equal(gx, g["gather_zero_dense"])
    np.testing.assert_array_equal(O.level1("scatter", idx, v, np.zeros(3))[2], g["scatter_into_zeros"])
    # alpha == 0 leaves the dense vector untouched even with NaN values
    out = O.level1("axpy", idx, np.full(2, np.nan), x, alpha=0.0)[2]
    np.testing.assert_array_equal(out, x)


def test_level1_dot_exact_integers():
    # integer data: every partial sum is exact, so the dot equals the integer
    # dot product whatever the association
    rng = np.random.default_rng(9)
    idx = np.sort(rng.choice(5000, 700, replace=False))
    v = rng.integers(-50, 50, 700).astype(np.float64)
    x = rng.integers(-50, 50, 5000).astype(np.float64)
    assert O.level1("dot", idx, v, x)[0] == float(np.dot(v.astype(np.int64), x[idx].astype(np.int64)))


def test_triangular_golden_vectors():
    g = GOLD["triangular"]
    vals = np.array(g["values"])
    x = O.trsv(0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], vals, 1.0, g["lower_n_rhs"])
    np.testing.assert_array_equal(x, g["lower_n_solution"])
    # CSC storage of the same lower matrix, conjugate transpose (values are the
    # CSC order of the same entries: (0,0)=2, (1,0)=1, (1,1)=3)
    x = O.trsv(1, 2, True, False, 2, g["csc_offsets"], g["csc_indices"], vals, 1.0, g["csc_lower_h_rhs"])
    np.testing.assert_array_equal(x, g["csc_lower_h_solution"])
    b = np.array(g["multiple_rhs_row_major"]).reshape(2, 2)
    x = O.trsv(0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], vals, 1.0, b)
    np.testing.assert_array_equal(x.reshape(-1), g["multiple_solution"])
    zv = np.array([float(t) for t in g["alpha_zero_values"]])
    x = O.trsv(0, 0, True, False, 2, g["csr_offsets"], g["csr_indices"], zv, 0.0, g["alpha_zero_rhs"])
    np.testing.assert_array_equal(x, g["alpha_zero_solution"])


def test_triangular_solves_recover_known_solutions():
    # build L with a dominant diagonal, pick x, form b = alpha^-1 * op(L) x in
    # exact integer arithmetic, and solve back
    rng = np.random.default_rng(2)
    n = 60
    dense = np.tril(rng.integers(-3, 4, (n, n)) * (rng.random((n, n)) < 0.2), -1).astype(np.float64)
    dense[np.arange(n), np.arange(n)] = 1.0  # unit-valued diagonal keeps every step exact
    off = np.concatenate([[0], np.cumsum((dense != 0).sum(1))])
    idx = np.concatenate([np.nonzero(r)[0] for r in dense])
    val = dense[dense != 0]
    x = rng.integers(-5, 6, n).astype(np.float64)
    for op, mat in ((0, dense), (1, dense.T)):
        b = mat @ x
        np.testing.assert_array_equal(O.trsv(0, op, True, False, n, off, idx, val, 1.0, b), x)
        np.testing.assert_array_equal(O.trsv(0, op, True, True, n, off, idx, val, 1.0, b), x)
