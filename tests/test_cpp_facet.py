"""The C++20 facet (include/asc_sparse_hip/hip.hpp, core_hip.hpp) through its
test program tests/cpp/facet_test.cc, in two builds:

- standalone (asc-cpp_amd/lib/facet_test): the facet's own restatement of
  the asc:: vocabulary;
- inside the reference's world (oracle/_ref/facet_in_tree, built by
  oracle/build_facet_in_tree.sh where the reference sources are present): the
  same facet source compiled with ASC_SPARSE_HIP_IN_ASC_TREE against the
  reference's own public headers and src/core/execution_internal.h, linked
  with the reference's core library compiled from its sources.  That build
  passing is the evidence that the facet drops into the reference tree with
  no call-site changes: its Status, Result, MemoryResource, Buffer,
  ExecutionContext (made by CreateHipExecutionContext over the reference's
  ExecutionState seam), CompletionEvent, CopyBytes and sparse views are the
  reference's.

Without a GPU the program checks host-view validation, the rejection matrix
on a validation-only context and the core provider's refusals; on a GPU it
also runs the reference's golden vectors (sparse_cuda_test.cc:240-341,
blas_test.cc:221-242, standard_blas_test.cc:176-334,
compressed_conversion_test.cc:26-34, coordinate_test.cc:96-142) through every
public facet function, uploads and downloads going through CopyBytes.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STANDALONE = os.path.join(ROOT, "asc-cpp_amd", "lib", "facet_test")
IN_TREE = os.path.join(ROOT, "oracle", "_ref", "facet_in_tree")
REFERENCE = "/root/reference"


def run(program, mode):
    assert os.path.exists(program), f"{program} is not built (run __graft_entry__.build())"
    p = subprocess.run([program, mode], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stderr
    assert "0 failures" in p.stdout


def in_tree_or_skip():
    if not os.path.exists(IN_TREE):
        pytest.skip("oracle/_ref/facet_in_tree is built only where the reference sources are")
    return IN_TREE


def test_facet_validation_without_gpu():
    run(STANDALONE, "validate")


def test_facet_in_reference_tree_validation_without_gpu():
    run(in_tree_or_skip(), "validate")


CONSUMER = r"""
#include "asc_sparse_hip/hip.hpp"
// A consumer written against the reference's CUDA facet, with Cuda -> Hip.
asc::Result<asc::CompletionEvent> Consume(asc::SparseHipContext& ctx, asc::MemoryResource& res,
                                          asc::CsrView<const double> host,
                                          const double* x, double* y, asc::MutableMemoryView ws) {
  auto clone = asc::HipCloneCsr(ctx, host, res);
  if (!clone.ok()) return clone.status();
  if (asc::Status s = clone->completion.Wait(); !s.ok()) return s;
  const auto& array = clone->array;
  auto view = array.view();
  if (!view.ok()) return view.status();
  auto xv = asc::HipStridedVectorView<const double>::Create(x, view->columns(), 1);
  auto yv = asc::HipStridedVectorView<double>::Create(y, view->rows(), 1);
  if (!xv.ok() || !yv.ok()) return asc::Status(asc::ErrorCode::kInvalidArgument, "views");
  auto need = asc::HipCsrSpmvWorkspaceSize(ctx, *view, *xv, *yv);
  if (!need.ok()) return need.status();
  return asc::HipCsrSpmv(ctx, 1.0, *view, *xv, 0.0, *yv, ws);
}
int main() {
  auto exec = asc::CreateHipExecutionContext(0, asc::Determinism::kDeterministic);
  if (!exec.ok()) return 1;
  auto ctx = asc::SparseHipContext::Create(std::move(*exec));
  return ctx.ok() ? 0 : 1;
}
"""


def test_facet_compiles_against_the_reference_headers(tmp_path):
    """A consumer TU, the facet header and its compiled sources, each checked
    with -fsyntax-only against the reference's own headers (the reference's
    types replace the facet's vocabulary)."""
    if not os.path.isdir(os.path.join(REFERENCE, "include", "asc")):
        pytest.skip("the reference tree is not present on this machine")
    cxx = shutil.which("g++")
    src = tmp_path / "consumer.cc"
    src.write_text(CONSUMER)
    flags = ["-std=c++20", "-fsyntax-only", "-Wall", "-DASC_SPARSE_HIP_IN_ASC_TREE",
             f"-I{REFERENCE}/include", f"-I{REFERENCE}/src/core", f"-I{ROOT}/include"]
    p = subprocess.run([cxx, *flags, str(src)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]
    hip_flags = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    for tu in ("core_hip.cc", "sparse_context.cc"):
        path = os.path.join(ROOT, "asc-cpp_amd", "csrc", "facet", tu)
        p = subprocess.run([cxx, *flags, *hip_flags, path], capture_output=True, text=True,
                           timeout=300)
        assert p.returncode == 0, (tu, p.stderr[-4000:])


@pytest.mark.gpu
def test_facet_golden_vectors_on_gpu():
    run(STANDALONE, "gpu")


@pytest.mark.gpu
def test_facet_in_reference_tree_golden_vectors_on_gpu():
    run(in_tree_or_skip(), "gpu")
