"""Host allocations on the C ABI's success paths: tests/cpp/alloc_test.cc
(built as asc-cpp_amd/lib/alloc_test) counts C++ heap allocations
process-wide while it repeats a warmed-up call -- compat SpMV (ordered,
merge, vector kernels, allocation-free entry point), standardized SpMV (op N
and op T with its device transpose), SpMM (plain and ordered) and the
Level-1 operations -- attributes each to the library or to the HIP runtime
by its call stack, and requires none from the library, as the reference
requires none of its serial Spmv and SparseDot
(tests/sparse/allocation_probe.cc:74-97, blas_test.cc:207-218,
standard_blas_test.cc:565-575)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "asc-cpp_amd", "lib", "alloc_test")


@pytest.mark.gpu
def test_success_paths_allocate_nothing_on_the_host():
    assert os.path.exists(PROGRAM), f"{PROGRAM} is not built (run __graft_entry__.build())"
    p = subprocess.run([PROGRAM, "gpu"], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout
