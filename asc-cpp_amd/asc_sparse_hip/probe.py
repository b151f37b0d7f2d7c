"""ctypes binding of the measurement probes (include/asc_sparse_hip_probe.h,
lib/libasc_sparse_probe.so): the ceilings the benchmark reports beside its
rooflines (x-element and B-row gather rates, HBM read and copy).  Not part of
the Sparse BLAS library."""
from __future__ import annotations

import ctypes as C
import os

from . import PROBE_PATH, Status, check, load

GATHER_PROBE_BLOCKS = 2048
ROW_GATHER_PROBE_BLOCKS = 8192
STREAM_PROBE_BLOCKS = 16384

_lib = None


def _load():
    global _lib
    if _lib is None:
        load()  # the library the probes link against
        if not os.path.exists(PROBE_PATH):
            raise ImportError(f"{PROBE_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(PROBE_PATH)
        vp = C.c_void_p
        lib.asc_sparse_hip_gather_probe.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int64, vp,
                                                    C.c_size_t]
        lib.asc_sparse_hip_gather_probe.restype = Status
        lib.asc_sparse_hip_row_gather_probe.argtypes = [vp, vp, C.c_int64, vp, C.c_int64, vp,
                                                        C.c_size_t]
        lib.asc_sparse_hip_row_gather_probe.restype = Status
        lib.asc_sparse_hip_stream_probe.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_size_t]
        lib.asc_sparse_hip_stream_probe.restype = Status
        _lib = lib
    return _lib


def _ptr(b):
    return b.ptr if hasattr(b, "ptr") else b


def gather_probe(ctx, indices, values, count: int, x, x_extent: int, partials):
    """Enqueues the gather-rate probe over device buffers."""
    check(_load().asc_sparse_hip_gather_probe(ctx.handle, _ptr(indices), _ptr(values), count,
                                              _ptr(x), x_extent, _ptr(partials),
                                              GATHER_PROBE_BLOCKS))


def row_gather_probe(ctx, indices, count: int, table, table_rows: int, partials):
    """Enqueues the 256-byte row gather probe over device buffers."""
    check(_load().asc_sparse_hip_row_gather_probe(ctx.handle, _ptr(indices), count, _ptr(table),
                                                  table_rows, _ptr(partials),
                                                  ROW_GATHER_PROBE_BLOCKS))


def stream_read(ctx, src, nbytes: int, partials):
    """Enqueues a plain HBM read of nbytes (word sums into partials)."""
    check(_load().asc_sparse_hip_stream_probe(ctx.handle, _ptr(src), None, nbytes,
                                              _ptr(partials), STREAM_PROBE_BLOCKS))


def stream_copy(ctx, src, dst, nbytes: int):
    """Enqueues a plain HBM copy of nbytes."""
    check(_load().asc_sparse_hip_stream_probe(ctx.handle, _ptr(src), _ptr(dst), nbytes, None, 0))
