"""ctypes binding of the seeded synthetic matrix generators (host side).

Families and value distributions mirror csrc/host/generators.h.  Every
workload of the benchmark and the tests is built here from a seed.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import GEN_PATH

UNIFORM_FIXED, ZIPF, PRECISION, STENCIL2D, STENCIL3D, BANDED, BLOCKDIAG8, BIPARTITE, \
    UNIFORM_VAR = range(9)
VALUES_UNIFORM, VALUES_NORMAL, VALUES_LOGNORMAL, VALUES_LAPLACIAN = range(4)

SEED_BASE = 0x5EEDA5C0  # + workload index


class Params(C.Structure):
    _fields_ = [("family", C.c_int32), ("value_dist", C.c_int32), ("rows", C.c_int64),
                ("cols", C.c_int64), ("per_row", C.c_int64), ("zipf_s", C.c_double),
                ("zipf_max", C.c_int64), ("band", C.c_int64), ("fill", C.c_double),
                ("grid", C.c_int64), ("seed", C.c_uint64), ("window", C.c_int64)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(GEN_PATH):
            raise ImportError(f"{GEN_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(GEN_PATH)
        lib.asc_gen_lengths.argtypes = [C.POINTER(Params), C.c_void_p, C.c_int]
        lib.asc_gen_lengths.restype = C.c_int64
        lib.asc_gen_fill.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_int]
        lib.asc_gen_fill.restype = C.c_int
        lib.asc_gen_vector.argtypes = [C.c_int64, C.c_uint64, C.c_int32, C.c_void_p, C.c_int]
        lib.asc_gen_vector.restype = None
        lib.asc_gen_zipf_mean.argtypes = [C.c_double, C.c_int64]
        lib.asc_gen_zipf_mean.restype = C.c_double
        _lib = lib
    return _lib


def threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


@dataclass
class Csr:
    rows: int
    cols: int
    offsets: np.ndarray  # int64[rows+1]
    indices: np.ndarray  # int32 or int64[nnz]
    values: np.ndarray   # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.values.shape[0])


def make(family: int, rows: int, cols: int | None = None, *, seed: int, values: int,
         per_row: int = 0, zipf_s: float = 1.5, zipf_max: int = 10_000, band: int = 0,
         fill: float = 0.5, grid: int = 0, window: int = 0, index_dtype=np.int32) -> Csr:
    cols = rows if cols is None else cols
    p = Params(family, values, rows, cols, per_row, zipf_s, zipf_max, band, fill, grid,
               seed & 0xFFFFFFFFFFFFFFFF, window)
    lib = _load()
    off = np.empty(rows + 1, dtype=np.int64)
    nnz = lib.asc_gen_lengths(C.byref(p), off.ctypes.data, threads())
    if nnz < 0:
        raise ValueError("invalid generator parameters")
    idx = np.empty(nnz, dtype=index_dtype)
    val = np.empty(nnz, dtype=np.float64)
    rc = lib.asc_gen_fill(C.byref(p), off.ctypes.data, idx.ctypes.data, idx.itemsize,
                          val.ctypes.data, threads())
    if rc != 0:
        raise ValueError("generator fill failed")
    return Csr(rows, cols, off, idx, val)


def vector(n: int, seed: int, values: int = VALUES_UNIFORM) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().asc_gen_vector(n, seed & 0xFFFFFFFFFFFFFFFF, values, out.ctypes.data, threads())
    return out


def zipf_mean(s: float, zmax: int) -> float:
    return float(_load().asc_gen_zipf_mean(s, zmax))


# ---- the benchmark workloads -------------------------------------------------

def uniform16(n: int = 4096, seed: int = SEED_BASE + 0) -> Csr:
    """Plain CSR SpMV case: n x n, 16 distinct sorted columns per row, U[-1,1]."""
    return make(UNIFORM_FIXED, n, n, seed=seed, values=VALUES_UNIFORM, per_row=16)


def zipf_large(n: int = 1 << 22, seed: int = SEED_BASE + 1) -> Csr:
    """Single large matrix: row lengths Zipf(1.5) clipped to [1, 1e4], N(0,1)."""
    return make(ZIPF, n, n, seed=seed, values=VALUES_NORMAL, zipf_s=1.5, zipf_max=10_000)


def precision_matrix(n: int = 1 << 20, seed: int = SEED_BASE + 3) -> Csr:
    """Precision case: 32 nnz/row (16 near-diagonal + 16 uniform), lognormal(0,2)."""
    return make(PRECISION, n, n, seed=seed, values=VALUES_LOGNORMAL)


def batched_specs(count: int = 10_000, seed: int = SEED_BASE + 2, lo: int = 64, hi: int = 65536):
    """Batched case: `count` matrices, rows log-uniform in [lo, hi], families
    cycling 2D 5-point Laplacian (k = round(sqrt(rows))), 3D 7-point
    Laplacian (k = round(cbrt(rows))) and random banded (half-bandwidth in
    U{8..64}, each in-band off-diagonal kept with p = 0.5, diagonal always
    kept, values U[-1,1])."""
    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    bands = rng.integers(8, 65, count)
    specs = []
    for i in range(count):
        fam = i % 3
        r = float(rows[i])
        if fam == 0:
            k = max(1, int(round(r ** 0.5)))
            specs.append(dict(family=STENCIL2D, rows=k * k, grid=k, values=VALUES_LAPLACIAN))
        elif fam == 1:
            k = max(1, int(round(r ** (1.0 / 3.0))))
            specs.append(dict(family=STENCIL3D, rows=k ** 3, grid=k, values=VALUES_LAPLACIAN))
        else:
            specs.append(dict(family=BANDED, rows=int(round(r)), band=int(bands[i]), fill=0.5,
                              values=VALUES_UNIFORM))
        specs[-1]["seed"] = (seed * 1_000_003 + i) & 0xFFFFFFFFFFFF
    return specs


def make_spec(spec: dict, index_dtype=np.int32) -> Csr:
    s = dict(spec)
    fam, rows = s.pop("family"), s.pop("rows")
    return make(fam, rows, rows, index_dtype=index_dtype, **s)


@dataclass
class Packed:
    """Many CSR matrices concatenated into shared arrays (host side)."""
    offsets: np.ndarray      # concatenated per-matrix offsets (each starts at 0)
    indices: np.ndarray
    values: np.ndarray
    rows: np.ndarray         # per-matrix rows
    nnz: np.ndarray          # per-matrix nnz
    off_start: np.ndarray    # element offset of each matrix's offsets
    nz_start: np.ndarray     # element offset of each matrix's indices/values
    vec_start: np.ndarray    # element offset of each matrix's x / y


def pack(mats: list) -> Packed:
    rows = np.array([m.rows for m in mats], dtype=np.int64)
    nnz = np.array([m.nnz for m in mats], dtype=np.int64)
    off_start = np.concatenate([[0], np.cumsum(rows + 1)[:-1]]).astype(np.int64)
    nz_start = np.concatenate([[0], np.cumsum(nnz)[:-1]]).astype(np.int64)
    vec_start = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    return Packed(np.concatenate([m.offsets for m in mats]),
                  np.concatenate([m.indices for m in mats]),
                  np.concatenate([m.values for m in mats]),
                  rows, nnz, off_start, nz_start, vec_start)
