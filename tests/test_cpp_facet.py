"""The C++20 facet (include/asc_sparse_hip/hip.hpp) through its test program
tests/cpp/facet_test.cc, built in-tree by the package Makefile.

Without a GPU the program checks host-view validation and the rejection
matrix on a validation-only context; on a GPU it also runs the reference's
golden vectors (sparse_cuda_test.cc:240-341, blas_test.cc:221-242,
standard_blas_test.cc:176-334, compressed_conversion_test.cc:26-34) through
every public facet function.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "asc-cpp_amd", "lib", "facet_test")


def run(mode):
    assert os.path.exists(PROGRAM), "facet_test is not built (run __graft_entry__.build())"
    p = subprocess.run([PROGRAM, mode], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stderr
    assert "0 failures" in p.stdout


def test_facet_validation_without_gpu():
    run("validate")


@pytest.mark.gpu
def test_facet_golden_vectors_on_gpu():
    run("gpu")
