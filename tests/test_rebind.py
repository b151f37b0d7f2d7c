"""RebindValues at the C ABI (compressed.h:472-516, coordinate.h:533-569;
sparse_cuda_test.cc:549-598): the descriptor keeps its structure, shape,
NNZ, memory space and trust; null, wrapping, structure-overlapping and
partly overlapping value spans are refused.  Host-only: nothing is
dereferenced, so fabricated addresses serve as device pointers.
"""
import ctypes as C

import pytest

import asc_sparse_hip as H


def desc(fmt=H.CSR, nnz=5, rows=3, cols=4, kind=H.F64, index_bytes=8):
    # offsets at 0x10000 (4 x 8 bytes), indices at 0x20000, values at 0x30000
    return H.SparseDesc(0x10000, 0x20000 if fmt != H.COO else None, 0x30000, rows, cols, nnz,
                        fmt, kind, index_bytes, 8, H.DEVICE, 1)


def rebind(d, ptr):
    out = H.SparseDesc()
    st = H.load().asc_sparse_hip_rebind_values(C.byref(d), ptr, C.byref(out))
    return st.code, out


def test_rebind_keeps_structure_and_trust():
    d = desc()
    code, out = rebind(d, 0x40000)
    assert code == 0
    assert (out.outer_offsets, out.inner_indices, out.values) == (0x10000, 0x20000, 0x40000)
    assert (out.rows, out.cols, out.nnz, out.format, out.element_kind, out.index_bytes,
            out.memory_space, out.trusted) == (3, 4, 5, H.CSR, H.F64, 8, H.DEVICE, 1)
    assert rebind(d, 0x30000)[0] == 0  # the same values


@pytest.mark.parametrize("ptr", [0x10000, 0x10000 + 24, 0x20000 - 8, 0x20000 + 32])
def test_rebind_rejects_structure_overlap(ptr):
    # offsets span [0x10000, 0x10020), indices [0x20000, 0x20028)
    assert rebind(desc(), ptr)[0] == H.INVALID_ARGUMENT


def test_rebind_boundaries_and_partial_overlap():
    assert rebind(desc(), 0x10020)[0] == 0          # just past the offsets
    assert rebind(desc(), 0x20028)[0] == 0          # just past the indices
    assert rebind(desc(), 0x30000 + 8)[0] == H.INVALID_ARGUMENT   # partial overlap
    assert rebind(desc(), 0x30000 - 8)[0] == H.INVALID_ARGUMENT
    assert rebind(desc(), 0x30000 + 40)[0] == 0      # adjacent, disjoint
    # narrowed indices span half the bytes: 0x20014 is free with 4-byte indices
    assert rebind(desc(index_bytes=4), 0x20014)[0] == 0
    assert rebind(desc(index_bytes=8), 0x20014)[0] == H.INVALID_ARGUMENT


def test_rebind_null_empty_and_wrap():
    assert rebind(desc(), None)[0] == H.INVALID_ARGUMENT
    assert rebind(desc(nnz=0), None)[0] == 0
    assert rebind(desc(), 2**64 - 16)[0] == H.OVERFLOW
    out = H.SparseDesc()
    assert H.load().asc_sparse_hip_rebind_values(None, 0x40000, C.byref(out)).code == H.INVALID_ARGUMENT
    assert H.load().asc_sparse_hip_rebind_values(C.byref(desc()), 0x40000, None).code == H.INVALID_ARGUMENT


def test_rebind_coordinate():
    d = desc(fmt=H.COO, nnz=2)  # coordinates span [0x10000, 0x10020)
    assert rebind(d, 0x10018)[0] == H.INVALID_ARGUMENT
    code, out = rebind(d, 0x10020)
    assert code == 0 and out.outer_offsets == 0x10000 and out.trusted == 1
