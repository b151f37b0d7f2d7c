"""Validation of the measurement probes' C ABI (libasc_sparse_probe.so) on a
validation-only context: every check that needs no GPU."""
from asc_sparse_hip import probe as P
import asc_sparse_hip as H


def test_gather_probe_validation(vctx):
    lib = P._load()
    probe = lambda idx, val, n, x, xn, out, m: lib.asc_sparse_hip_gather_probe(
        vctx.handle, idx, val, n, x, xn, out, m).code
    i, v, x, out = 0x100000, 0x200000, 0x300000, 0x400000
    assert probe(i, v, 10, x, 100, out, P.GATHER_PROBE_BLOCKS) == H.UNAVAILABLE  # valid
    assert probe(i, v, -1, x, 100, out, P.GATHER_PROBE_BLOCKS) == H.SHAPE
    assert probe(i, v, 10, x, 0, out, P.GATHER_PROBE_BLOCKS) == H.SHAPE
    assert probe(i, v, 10, x, 2 ** 33, out, P.GATHER_PROBE_BLOCKS) == H.UNSUPPORTED
    assert probe(i, v, 10, x, 100, out, P.GATHER_PROBE_BLOCKS - 1) == H.INVALID_ARGUMENT
    assert probe(i, v, 10, x, 100, None, P.GATHER_PROBE_BLOCKS) == H.INVALID_ARGUMENT
    assert probe(i, v, 10, x, 100, v + 8, P.GATHER_PROBE_BLOCKS) == H.INVALID_ARGUMENT  # overlaps values
    assert probe(None, None, 0, None, 0, out, P.GATHER_PROBE_BLOCKS) == H.UNAVAILABLE  # empty is fine


def test_stream_probe_validation(vctx):
    lib = P._load()
    probe = lambda src, dst, n, out, m: lib.asc_sparse_hip_stream_probe(
        vctx.handle, src, dst, n, out, m).code
    src, dst, out = 0x100000, 0x900000, 0x1100000
    assert probe(src, None, 1 << 16, out, P.STREAM_PROBE_BLOCKS) == H.UNAVAILABLE  # valid read
    assert probe(src, dst, 1 << 16, None, 0) == H.UNAVAILABLE  # valid copy
    assert probe(src, None, 100, out, P.STREAM_PROBE_BLOCKS) == H.SHAPE  # not a multiple of 16
    assert probe(src, None, -16, out, P.STREAM_PROBE_BLOCKS) == H.SHAPE
    assert probe(src + 8, None, 1 << 16, out, P.STREAM_PROBE_BLOCKS) == H.INVALID_ARGUMENT
    assert probe(src, None, 1 << 16, out, P.STREAM_PROBE_BLOCKS - 1) == H.INVALID_ARGUMENT
    assert probe(src, src + 1024, 1 << 16, None, 0) == H.INVALID_ARGUMENT  # overlap
    assert probe(src, None, 1 << 16, src + 16, P.STREAM_PROBE_BLOCKS) == H.INVALID_ARGUMENT
    assert probe(None, None, 0, None, 0) == H.UNAVAILABLE  # empty is fine


def test_row_gather_probe_validation(vctx):
    lib = P._load()
    probe = lambda idx, n, t, rows, out, m: lib.asc_sparse_hip_row_gather_probe(
        vctx.handle, idx, n, t, rows, out, m).code
    i, t, out = 0x100000, 0x200000, 0x900000
    B = P.ROW_GATHER_PROBE_BLOCKS
    assert probe(i, 10, t, 100, out, B) == H.UNAVAILABLE  # valid
    assert probe(i, -1, t, 100, out, B) == H.SHAPE
    assert probe(i, 10, t, 0, out, B) == H.SHAPE
    assert probe(i, 10, t, 2 ** 33, out, B) == H.UNSUPPORTED
    assert probe(i, 10, t, 100, out, B - 1) == H.INVALID_ARGUMENT
    assert probe(i, 10, t, 100, None, B) == H.INVALID_ARGUMENT
    assert probe(i, 10, t + 8, 100, out, B) == H.INVALID_ARGUMENT  # table not 32-byte aligned
    assert probe(i, 10, t, 100, t + 256, B) == H.INVALID_ARGUMENT  # partials overlap the table
    assert probe(None, 0, None, 0, out, B) == H.UNAVAILABLE  # empty is fine
