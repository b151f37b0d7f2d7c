// Tests of the C++20 facet (include/asc_sparse_hip/hip.hpp) and its core
// provider (core_hip.hpp), written the way the reference tests its CUDA facet
// (tests/sparse_cuda/sparse_cuda_test.cc): golden vectors restated from the
// reference's tests, element types float and double, and the rejection
// matrix.  The same source builds standalone (the facet's own vocabulary) and
// inside the asc-cpp tree (ASC_SPARSE_HIP_IN_ASC_TREE: the reference's core and
// sparse headers and its compiled core library), so it checks that the facet
// is a drop-in on the reference's types: uploads go through the reference's
// CopyBytes on a HIP ExecutionContext, arrays come from a MemoryResource&.
//
//   facet_test validate   no GPU needed: host-view validation and every
//                         rejection on a validation-only context
//   facet_test gpu        the golden vectors on device 0
//
// Exit status 0 when every check passes; each failing check prints a line.
#include <cmath>
#include <cstdint>
#include <memory>
#include <utility>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "asc_sparse_hip/hip.hpp"

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                              \
  do {                                                                           \
    ++g_checks;                                                                  \
    if (!(cond)) {                                                               \
      ++g_failures;                                                              \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
    }                                                                            \
  } while (0)

#define CHECK_CODE(status_expr, expected)                                        \
  do {                                                                           \
    const asc::Status _s = (status_expr);                                        \
    ++g_checks;                                                                  \
    if (_s.code() != (expected)) {                                               \
      ++g_failures;                                                              \
      std::fprintf(stderr, "FAILED %s:%d: code %d (%s), expected %d\n", __FILE__, \
                   __LINE__, static_cast<int>(_s.code()), _s.message().c_str(),  \
                   static_cast<int>(expected));                                  \
    }                                                                            \
  } while (0)

void Case(const char* name) {
  std::printf("case %s\n", name);
  std::fflush(stdout);
}

using asc::ErrorCode;
using asc::MemorySpace;
using Csr = asc::SparseCompressedFormat;

asc::MutableMemoryView NoWorkspace() { return asc::MutableMemoryView(nullptr, 0, MemorySpace::kDevice); }

// tests/sparse/blas_test.cc:123-125 and sparse_cuda_test.cc:240-341.
const std::vector<asc::nnz_t> kOffsets3x4{0, 2, 4, 5};
const std::vector<asc::index_t> kIndices3x4{1, 3, 0, 2, 1};
const std::vector<double> kValues3x4{2, -1, 4, 5, 3};

const std::vector<asc::nnz_t> kOffsets4x5{0, 2, 3, 6, 8};
const std::vector<asc::index_t> kIndices4x5{0, 3, 1, 0, 2, 4, 1, 4};
const std::vector<double> kValues4x5{2, -1, 4, 7, 3, 5, -2, 6};

template <typename E>
std::vector<E> Cast(const std::vector<double>& v) {
  return std::vector<E>(v.begin(), v.end());
}

// ---- validation (no GPU) ---------------------------------------------------------

void TestHostViewValidation() {
  Case("host view validation");
  const std::array<asc::extent_t, 2> shape{3, 4};
  auto ok = asc::CsrView<const double>::Create(kOffsets3x4.data(), kIndices3x4.data(),
                                              kValues3x4.data(), shape, 5, MemorySpace::kHost);
  CHECK(ok.ok());
  // compressed.h:711-746: unsorted and out-of-range indices, wrong final offset
  const std::vector<asc::index_t> unsorted{3, 1, 0, 2, 1};
  CHECK_CODE(asc::CsrView<const double>::Create(kOffsets3x4.data(), unsorted.data(),
                                                kValues3x4.data(), shape, 5, MemorySpace::kHost)
                 .status(),
             ErrorCode::kInvalidArgument);
  const std::vector<asc::index_t> out_of_range{1, 4, 0, 2, 1};
  CHECK_CODE(asc::CsrView<const double>::Create(kOffsets3x4.data(), out_of_range.data(),
                                                kValues3x4.data(), shape, 5, MemorySpace::kHost)
                 .status(),
             ErrorCode::kIndex);
  CHECK_CODE(asc::CsrView<const double>::Create(kOffsets3x4.data(), kIndices3x4.data(),
                                                kValues3x4.data(), shape, 4, MemorySpace::kHost)
                 .status(),
             ErrorCode::kInvalidArgument);
  const std::array<asc::extent_t, 2> tiny{1, 1};
  CHECK_CODE(asc::CsrView<const double>::Create(kOffsets3x4.data(), kIndices3x4.data(),
                                                kValues3x4.data(), tiny, 5, MemorySpace::kHost)
                 .status(),
             ErrorCode::kShape);
  // RebindValues (compressed.h:472-516; sparse_cuda_test.cc:549-598): the
  // same pointer and a distinct buffer rebind; structure and partial overlap
  // are refused.
  CHECK(ok->RebindValues(kValues3x4.data()).ok());
  const std::vector<double> other{9, 8, 7, 6, 5, 4};
  auto rebound = ok->RebindValues(other.data());
  CHECK(rebound.ok() && rebound->values() == other.data() &&
        rebound->outer_offsets() == kOffsets3x4.data() && rebound->nnz() == 5 &&
        rebound->canonical_structure_trusted() == ok->canonical_structure_trusted());
  CHECK_CODE(ok->RebindValues(reinterpret_cast<const double*>(kOffsets3x4.data())).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(ok->RebindValues(reinterpret_cast<const double*>(kIndices3x4.data())).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(rebound->RebindValues(other.data() + 1).status(), ErrorCode::kInvalidArgument);
  CHECK_CODE(ok->RebindValues(static_cast<const double*>(nullptr)).status(),
             ErrorCode::kInvalidArgument);
  const std::vector<asc::index_t> coords{0, 1, 2, 3};
  const std::array<asc::extent_t, 2> cshape{3, 4};
  auto coo = asc::CoordinateView<const double, 2>::Create(coords.data(), other.data(), cshape, 2,
                                                          MemorySpace::kHost);
  CHECK(coo.ok());
  CHECK(coo->RebindValues(other.data() + 2).ok());
  CHECK_CODE(coo->RebindValues(other.data() + 1).status(), ErrorCode::kInvalidArgument);
  CHECK_CODE(coo->RebindValues(reinterpret_cast<const double*>(coords.data() + 1)).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(asc::HipStridedVectorView<double>::Create(nullptr, 4, 0).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(asc::HipStridedVectorView<double>::Create(nullptr, -1, 1).status(),
             ErrorCode::kInvalidArgument);
}

void TestRejectionMatrix() {
  Case("rejection matrix on a validation-only context");
  auto ctx = asc::SparseHipContext::CreateValidationOnly();
  CHECK(ctx.ok());
  if (!ctx.ok()) return;
  // A device view made by Create is untrusted (compressed.h:759-770).
  const std::array<asc::extent_t, 2> shape{3, 4};
  auto* fake = reinterpret_cast<double*>(std::uintptr_t{0x10000});
  auto untrusted = asc::CsrView<const double>::Create(
      reinterpret_cast<const asc::nnz_t*>(0x20000), reinterpret_cast<const asc::index_t*>(0x30000),
      fake, shape, 5, MemorySpace::kDevice);
  CHECK(untrusted.ok());
  auto x = asc::HipStridedVectorView<const double>::Create(reinterpret_cast<const double*>(0x40000), 4, 1);
  auto y = asc::HipStridedVectorView<double>::Create(reinterpret_cast<double*>(0x50000), 3, 1);
  CHECK(x.ok() && y.ok());
  CHECK_CODE(asc::HipCsrSpmvWorkspaceSize(*ctx, *untrusted, *x, *y).status(),
             ErrorCode::kInvalidArgument);
  // A trusted device view reaches the enqueue point: kUnavailable.
  const auto trusted =
      asc::internal_sparse_compressed::ProviderAccess::MakeCanonicalView<const double, Csr::kCsr>(
          reinterpret_cast<const asc::nnz_t*>(0x20000), reinterpret_cast<const asc::index_t*>(0x30000),
          fake, {3, 4}, 5, MemorySpace::kDevice);
  auto ws = asc::HipCsrSpmvWorkspaceSize(*ctx, trusted, *x, *y);
  CHECK(ws.ok());
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, trusted, *x, 0.0, *y, NoWorkspace(), asc::HipSpmvAlgorithm::kVector8).status(),
             ErrorCode::kUnavailable);
  // Merge-path needs its workspace.
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, trusted, *x, 0.0, *y, NoWorkspace(), asc::HipSpmvAlgorithm::kMergePath).status(),
             ErrorCode::kInvalidArgument);
  // Shape mismatches (blas.h:140-205).
  auto x5 = asc::HipStridedVectorView<const double>::Create(reinterpret_cast<const double*>(0x40000), 5, 1);
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, trusted, *x5, 0.0, *y, NoWorkspace()).status(), ErrorCode::kShape);
  // Output overlapping the matrix values.
  auto y_on_values = asc::HipStridedVectorView<double>::Create(fake, 3, 1);
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, trusted, *x, 0.0, *y_on_values, NoWorkspace()).status(),
             ErrorCode::kInvalidArgument);
  // Host matrices are not device operands.
  auto host = asc::CsrView<const double>::Create(kOffsets3x4.data(), kIndices3x4.data(),
                                                kValues3x4.data(), shape, 5, MemorySpace::kHost);
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, *host, *x, 0.0, *y, NoWorkspace()).status(), ErrorCode::kMemoryAccess);
  // Column-panel algorithm without an analysis.
  CHECK_CODE(asc::HipCsrSpmv(*ctx, 1.0, trusted, *x, 0.0, *y, NoWorkspace(),
                             asc::HipSpmvAlgorithm::kColumnPanel)
                 .status(),
             ErrorCode::kInvalidArgument);
  // Standardized SpMV with a wrong output extent.
  auto vx = asc::SparseBlasVectorView<const double>::Create(
      reinterpret_cast<const double*>(0x40000), 4, 1,
      asc::ConstMemoryView(reinterpret_cast<const void*>(0x40000), 32, MemorySpace::kDevice));
  auto vy = asc::SparseBlasVectorView<double>::Create(
      reinterpret_cast<double*>(0x50000), 4, 1,
      asc::ConstMemoryView(reinterpret_cast<const void*>(0x50000), 32, MemorySpace::kDevice));
  CHECK_CODE(asc::HipSpmv(*ctx, asc::SparseBlasTranspose::kNone, 1.0, trusted, *vx, *vy).status(),
             ErrorCode::kShape);
  // Core provider: a HIP context comes only from its factory, and the
  // sparse context takes only a HIP execution context.
  CHECK_CODE(asc::ExecutionContext::Create(asc::Backend::kHip, asc::Device{asc::Backend::kHip, 0})
                 .status(),
             ErrorCode::kUnsupported);
  CHECK_CODE(asc::CreateHipExecutionContext(-1).status(), ErrorCode::kInvalidArgument);
  CHECK_CODE(asc::CreateHipExecutionContext(1 << 20).status(), ErrorCode::kUnavailable);
  CHECK_CODE(asc::SparseHipContext::Create(asc::ExecutionContext::Serial()).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(asc::RecordHipEvent(asc::ExecutionContext::Serial()).status(),
             ErrorCode::kInvalidArgument);
  CHECK_CODE(asc::HipMemoryResource::Create(0, MemorySpace::kHost).status(),
             ErrorCode::kInvalidArgument);
}

// ---- GPU ---------------------------------------------------------------------------

struct Gpu {
  asc::SparseHipContext ctx;
  std::unique_ptr<asc::HipMemoryResource> device;
};

asc::Buffer Allocate(Gpu& g, std::size_t bytes) {
  auto b = asc::Buffer::Allocate(*g.device, bytes, 256);
  CHECK(b.ok());
  return b.ok() ? std::move(*b) : asc::Buffer(std::move(*asc::Buffer::Allocate(*g.device, 0)));
}

asc::MutableMemoryView Whole(asc::Buffer& b) {
  return asc::MutableMemoryView(b.data(), b.size(), MemorySpace::kDevice);
}

// Host <-> device through the reference's CopyBytes on the HIP context.
bool Copy(Gpu& g, void* dst, MemorySpace dst_space, const void* src, MemorySpace src_space,
          std::size_t bytes) {
  auto e = asc::CopyBytes(g.ctx.execution_context(), asc::MutableMemoryView(dst, bytes, dst_space),
                          asc::ConstMemoryView(src, bytes, src_space), bytes);
  return e.ok() && e->Wait().ok();
}
bool FromDevice(Gpu& g, void* dst, const void* src, std::size_t bytes) {
  return Copy(g, dst, MemorySpace::kHost, src, MemorySpace::kDevice, bytes);
}

template <typename E>
asc::Buffer Upload(Gpu& g, const std::vector<E>& host) {
  asc::Buffer b = Allocate(g, sizeof(E) * host.size() + 8);
  if (!host.empty())
    CHECK(Copy(g, b.data(), MemorySpace::kDevice, host.data(), MemorySpace::kHost,
               sizeof(E) * host.size()));
  return b;
}

template <typename E>
std::vector<E> Download(Gpu& g, const asc::Buffer& b, std::size_t n) {
  std::vector<E> host(n);
  if (n) CHECK(FromDevice(g, host.data(), b.data(), sizeof(E) * n));
  return host;
}

template <typename E>
asc::HipCsrArray<E> CloneCsr(Gpu& g, const std::vector<asc::nnz_t>& off, const std::vector<asc::index_t>& idx,
                             const std::vector<E>& val, asc::extent_t rows, asc::extent_t cols,
                             asc::HipIndexWidth width = asc::HipIndexWidth::k64) {
  const std::array<asc::extent_t, 2> shape{rows, cols};
  auto host = asc::CsrView<const E>::Create(off.data(), idx.data(), val.data(), shape,
                                            static_cast<asc::nnz_t>(idx.size()), MemorySpace::kHost);
  CHECK(host.ok());
  auto clone = asc::HipCloneCsr(g.ctx, *host, *g.device, width);
  CHECK(clone.ok());
  CHECK(clone->completion.Wait().ok());
  return std::move(clone->array);
}

const asc::HipSpmvAlgorithm kAlgorithms[] = {
    asc::HipSpmvAlgorithm::kOrdered,  asc::HipSpmvAlgorithm::kMergePath,
    asc::HipSpmvAlgorithm::kVector1,  asc::HipSpmvAlgorithm::kVector4,
    asc::HipSpmvAlgorithm::kVector64, asc::HipSpmvAlgorithm::kAuto};

// sparse_cuda_test.cc:240-341: strided device golden, alpha=2, beta=-0.5 and
// alpha=1, beta=0 with NaN in y.
// The HIP core provider on the reference's seams: CopyBytes round trip
// through pinned and device memory, events, a Buffer's release.
void TestCoreProvider(Gpu& g) {
  Case("core provider: CopyBytes, events, resources");
  auto pinned = asc::HipMemoryResource::Create(0, MemorySpace::kPinnedHost);
  CHECK(pinned.ok());
  if (!pinned.ok()) return;
  auto p = asc::Buffer::Allocate(**pinned, 64, 64);
  CHECK(p.ok() && p->valid() && p->size() == 64);
  const std::vector<double> src{1.5, -2.25, 3.0, 4.0e300, -0.0, 6, 7, 8};
  std::vector<double> back(8, 0.0);
  asc::Buffer d = Allocate(g, 64);
  CHECK(Copy(g, p->data(), MemorySpace::kPinnedHost, src.data(), MemorySpace::kHost, 64));
  CHECK(Copy(g, d.data(), MemorySpace::kDevice, p->data(), MemorySpace::kPinnedHost, 64));
  CHECK(FromDevice(g, back.data(), d.data(), 64));
  CHECK(std::memcmp(back.data(), src.data(), 64) == 0);
  // A device view declared as host memory is refused.
  CHECK_CODE(asc::CopyBytes(g.ctx.execution_context(),
                            asc::MutableMemoryView(back.data(), 64, MemorySpace::kHost),
                            asc::ConstMemoryView(d.data(), 64, MemorySpace::kHost), 64)
                 .status(),
             ErrorCode::kMemoryAccess);
  auto ev = asc::RecordHipEvent(g.ctx.execution_context());
  CHECK(ev.ok() && ev->valid() && ev->Wait().ok() && ev->is_complete());
  auto q = ev->Query();
  CHECK(q.ok() && *q);
  CHECK(g.ctx.execution_context().CanAccess(MemorySpace::kDevice));
  CHECK(g.ctx.Synchronize().ok());
}

template <typename E, typename Matrix>
void SpmvGoldenOn(Gpu& g, const Matrix& cv) {
  auto xb = Upload<E>(g, {1, -2, 3, 4, -1});
  for (auto alg : kAlgorithms) {
    auto yb = Upload<E>(g, {10, 20, 30, 40});
    auto x = asc::HipStridedVectorView<const E>::Create(static_cast<const E*>(xb.data()), 5, 1);
    auto y = asc::HipStridedVectorView<E>::Create(static_cast<E*>(yb.data()), 4, 1);
    auto need = asc::HipCsrSpmvWorkspaceSize(g.ctx, cv, *x, *y, alg);
    CHECK(need.ok());
    asc::Buffer ws = Allocate(g, *need + 256);
    auto done = asc::HipCsrSpmv(g.ctx, E(2), cv, *x, E(-0.5), *y, Whole(ws), alg);
    CHECK(done.ok() && done->Wait().ok());
    const auto out = Download<E>(g, yb, 4);
    CHECK(out[0] == E(-9) && out[1] == E(-26) && out[2] == E(7) && out[3] == E(-24));
    auto nanb = Upload<E>(g, {NAN, NAN, NAN, NAN});
    auto yn = asc::HipStridedVectorView<E>::Create(static_cast<E*>(nanb.data()), 4, 1);
    done = asc::HipCsrSpmv(g.ctx, E(1), cv, *x, E(0), *yn, Whole(ws), alg);
    CHECK(done.ok() && done->Wait().ok());
    const auto out2 = Download<E>(g, nanb, 4);
    CHECK(out2[0] == E(-2) && out2[1] == E(-8) && out2[2] == E(11) && out2[3] == E(-2));
  }
}

template <typename E>
void TestCsrSpmvGolden(Gpu& g) {
  Case(sizeof(E) == 8 ? "compat SpMV golden (double)" : "compat SpMV golden (float)");
  auto a = CloneCsr<E>(g, kOffsets4x5, kIndices4x5, Cast<E>(kValues4x5), 4, 5);
  const auto& ca = a;
  auto view = ca.view();
  CHECK(view.ok() && view->canonical_structure_trusted());
  CHECK_CODE(ca.narrow_view().status(), ErrorCode::kInvalidState);
  SpmvGoldenOn<E>(g, *view);
  // The explicit 32-bit-index clone, through its facet-side view.
  auto n = CloneCsr<E>(g, kOffsets4x5, kIndices4x5, Cast<E>(kValues4x5), 4, 5, asc::HipIndexWidth::k32);
  const auto& cn = n;
  CHECK_CODE(cn.view().status(), ErrorCode::kInvalidState);
  auto nview = cn.narrow_view();
  CHECK(nview.ok());
  SpmvGoldenOn<E>(g, *nview);
}

// blas_test.cc:221-242: [1e20, 1, -1e20] . [1, 1, 1] is exactly 0 in float
// under the reference's left-to-right order.
void TestSummationOrder(Gpu& g) {
  Case("summation order (float, ordered)");
  auto a = CloneCsr<float>(g, {0, 3}, {0, 1, 2}, {1e20f, 1.0f, -1e20f}, 1, 3);
  const asc::CsrView<const float> cv = *std::as_const(a).view();
  auto xb = Upload<float>(g, {1, 1, 1});
  auto yb = Upload<float>(g, {-1});
  auto x = asc::HipStridedVectorView<const float>::Create(static_cast<const float*>(xb.data()), 3, 1);
  auto y = asc::HipStridedVectorView<float>::Create(static_cast<float*>(yb.data()), 1, 1);
  auto done = asc::HipCsrSpmv(g.ctx, 1.0f, cv, *x, 0.0f, *y, NoWorkspace(), asc::HipSpmvAlgorithm::kOrdered);
  CHECK(done.ok() && done->Wait().ok());
  CHECK(Download<float>(g, yb, 1)[0] == 0.0f);
}

// standard_blas_test.cc:176-334 (real parts): 2x2 A = [[2,0],[1,3]].
template <typename E>
void TestStandard(Gpu& g) {
  Case(sizeof(E) == 8 ? "standardized SpMV/SpMM golden (double)" : "standardized SpMV/SpMM golden (float)");
  auto a = CloneCsr<E>(g, {0, 1, 3}, {0, 0, 1}, {2, 1, 3}, 2, 2);
  const asc::CsrView<const E> cv = *std::as_const(a).view();
  auto xb = Upload<E>(g, {1, -1});
  const asc::ConstMemoryView xs(xb.data(), 2 * sizeof(E), MemorySpace::kDevice);
  auto x = asc::SparseBlasVectorView<const E>::Create(static_cast<const E*>(xb.data()), 2, 1, xs);
  for (auto [op, expect] : {std::pair{asc::SparseBlasTranspose::kNone, std::array<E, 2>{2, -2}},
                            std::pair{asc::SparseBlasTranspose::kTranspose, std::array<E, 2>{1, -3}}}) {
    auto yb = Upload<E>(g, {0, 0});
    const asc::ConstMemoryView ys(yb.data(), 2 * sizeof(E), MemorySpace::kDevice);
    auto y = asc::SparseBlasVectorView<E>::Create(static_cast<E*>(yb.data()), 2, 1, ys);
    auto need = asc::HipSpmvWorkspaceSize(g.ctx, op, cv, *x, *y);
    CHECK(need.ok());
    asc::Buffer ws = Allocate(g, *need + 256);
    auto done = asc::HipSpmv(g.ctx, op, E(1), cv, *x, *y, Whole(ws));
    CHECK(done.ok() && done->Wait().ok());
    const auto out = Download<E>(g, yb, 2);
    CHECK(out[0] == expect[0] && out[1] == expect[1]);
  }
  // Row-major SpMM: C += A*B with B = [[1,2],[3,4]] gives [[2,4],[10,14]].
  auto bb = Upload<E>(g, {1, 2, 3, 4});
  auto cb = Upload<E>(g, {0, 0, 0, 0});
  const asc::ConstMemoryView bs(bb.data(), 4 * sizeof(E), MemorySpace::kDevice);
  const asc::ConstMemoryView cs(cb.data(), 4 * sizeof(E), MemorySpace::kDevice);
  auto b = asc::SparseBlasMatrixView<const E>::Create(static_cast<const E*>(bb.data()), 2, 2,
                                                      asc::SparseBlasLayout::kRowMajor, 2, bs);
  auto c = asc::SparseBlasMatrixView<E>::Create(static_cast<E*>(cb.data()), 2, 2,
                                                asc::SparseBlasLayout::kRowMajor, 2, cs);
  CHECK(b.ok() && c.ok());
  auto done = asc::HipSpmm(g.ctx, asc::SparseBlasTranspose::kNone, E(1), cv, *b, *c);
  CHECK(done.ok() && done->Wait().ok());
  const auto out = Download<E>(g, cb, 4);
  CHECK(out[0] == 2 && out[1] == 4 && out[2] == 10 && out[3] == 14);
  // alpha == 0 is a no-op: C keeps its values (standard_blas_test.cc:317-334).
  done = asc::HipSpmm(g.ctx, asc::SparseBlasTranspose::kNone, E(0), cv, *b, *c);
  CHECK(done.ok() && done->Wait().ok());
  CHECK(Download<E>(g, cb, 4) == out);
  // The ordered form adds the same exact products again: [[4,8],[20,28]].
  done = asc::HipSpmm(g.ctx, asc::SparseBlasTranspose::kNone, E(1), cv, *b, *c, NoWorkspace(),
                      asc::HipSpmmForm::kOrdered);
  CHECK(done.ok() && done->Wait().ok());
  const auto twice = Download<E>(g, cb, 4);
  CHECK(twice[0] == 4 && twice[1] == 8 && twice[2] == 20 && twice[3] == 28);
  // A 4-column row-major block takes the plain kernel: B = [[1..4],[5..8]],
  // C = A*B = [[2,4,6,8],[16,20,24,28]].
  auto b4b = Upload<E>(g, {1, 2, 3, 4, 5, 6, 7, 8});
  auto c4b = Upload<E>(g, {0, 0, 0, 0, 0, 0, 0, 0});
  const asc::ConstMemoryView b4s(b4b.data(), 8 * sizeof(E), MemorySpace::kDevice);
  const asc::ConstMemoryView c4s(c4b.data(), 8 * sizeof(E), MemorySpace::kDevice);
  auto b4 = asc::SparseBlasMatrixView<const E>::Create(static_cast<const E*>(b4b.data()), 2, 4,
                                                       asc::SparseBlasLayout::kRowMajor, 4, b4s);
  auto c4 = asc::SparseBlasMatrixView<E>::Create(static_cast<E*>(c4b.data()), 2, 4,
                                                 asc::SparseBlasLayout::kRowMajor, 4, c4s);
  CHECK(b4.ok() && c4.ok());
  auto need4 = asc::HipSpmmWorkspaceSize(g.ctx, asc::SparseBlasTranspose::kNone, cv, *b4, *c4);
  CHECK(need4.ok());
  done = asc::HipSpmm(g.ctx, asc::SparseBlasTranspose::kNone, E(1), cv, *b4, *c4);
  CHECK(done.ok() && done->Wait().ok());
  const auto out4 = Download<E>(g, c4b, 8);
  CHECK(out4 == (std::vector<E>{2, 4, 6, 8, 16, 20, 24, 28}));
}

// compressed_conversion_test.cc:26-34,305-400: CSR <-> CSC <-> COO, bit-exact.
template <typename E>
void TestConversions(Gpu& g) {
  Case(sizeof(E) == 8 ? "conversions golden (double)" : "conversions golden (float)");
  auto a = CloneCsr<E>(g, kOffsets3x4, kIndices3x4, Cast<E>(kValues3x4), 3, 4);
  const asc::CsrView<const E> cv = *std::as_const(a).view();
  auto need = asc::HipTransposeWorkspaceSize(g.ctx, cv);
  CHECK(need.ok());
  asc::Buffer ws = Allocate(g, *need + 256);
  auto csc = asc::HipConvertToCsc(g.ctx, cv, *g.device, Whole(ws));
  CHECK(csc.ok() && csc->completion.Wait().ok());
  {
    auto view = std::as_const(csc->array).view();
    CHECK(view.ok());
    std::vector<asc::nnz_t> o(5);
    std::vector<asc::index_t> i(5);
    std::vector<E> v(5);
    CHECK(FromDevice(g, o.data(), view->outer_offsets(), 5 * 8));
    CHECK(FromDevice(g, i.data(), view->inner_indices(), 5 * 8));
    CHECK(FromDevice(g, v.data(), view->values(), 5 * sizeof(E)));
    CHECK((o == std::vector<asc::nnz_t>{0, 1, 3, 4, 5}));
    CHECK((i == std::vector<asc::index_t>{1, 0, 2, 1, 0}));
    CHECK((v == std::vector<E>{4, 2, 3, 5, -1}));
    // CSC -> CSR returns the original arrays.
    const asc::CscView<const E> ccv = *view;
    auto back = asc::HipConvertToCsr(g.ctx, ccv, *g.device, Whole(ws));
    CHECK(back.ok() && back->completion.Wait().ok());
    auto bv = std::as_const(back->array).view();
    std::vector<asc::nnz_t> bo(4);
    std::vector<asc::index_t> bi(5);
    std::vector<E> bvals(5);
    CHECK(FromDevice(g, bo.data(), bv->outer_offsets(), 4 * 8));
    CHECK(FromDevice(g, bi.data(), bv->inner_indices(), 5 * 8));
    CHECK(FromDevice(g, bvals.data(), bv->values(), 5 * sizeof(E)));
    CHECK(bo == kOffsets3x4);
    CHECK(bi == kIndices3x4);
    CHECK(bvals == Cast<E>(kValues3x4));
  }
  // CSR -> COO -> CSR.
  auto coo = asc::HipConvertToCoordinate(g.ctx, cv, *g.device);
  CHECK(coo.ok() && coo->completion.Wait().ok());
  auto coov = coo->array.view();
  CHECK(coov.ok());
  std::vector<asc::index_t> pairs(10);
  CHECK(FromDevice(g, pairs.data(), coov->coordinates(), 10 * 8));
  CHECK((pairs == std::vector<asc::index_t>{0, 1, 0, 3, 1, 0, 1, 2, 2, 1}));
  auto cneed = asc::HipConvertCoordinateWorkspaceSize(g.ctx, *coov, asc::SparseCompressedFormat::kCsr);
  CHECK(cneed.ok());
  asc::Buffer cws = Allocate(g, *cneed + 256);
  auto csr = asc::HipConvertToCsr(g.ctx, *coov, *g.device, Whole(cws));
  CHECK(csr.ok() && csr->completion.Wait().ok());
  auto rv = std::as_const(csr->array).view();
  std::vector<asc::nnz_t> ro(4);
  std::vector<asc::index_t> ri(5);
  CHECK(FromDevice(g, ro.data(), rv->outer_offsets(), 4 * 8));
  CHECK(FromDevice(g, ri.data(), rv->inner_indices(), 5 * 8));
  CHECK(ro == kOffsets3x4);
  CHECK(ri == kIndices3x4);
}

// Column-panel analysis on the golden matrix: panels of 1 column, every row
// listed (long_row_min 1), with the default and the stream staging; the
// result still meets the golden vector.
void TestColumnPanel(Gpu& g, uint8_t staging) {
  Case(staging == ASC_HIP_STAGING_STREAM ? "column-panel plan, stream staging (double)"
                                         : "column-panel plan (double)");
  auto a = CloneCsr<double>(g, kOffsets4x5, kIndices4x5, kValues4x5, 4, 5);
  const asc::CsrView<const double> cv = *std::as_const(a).view();
  asc_hip_panel_options opt{1, 1, ASC_HIP_ALG_AUTO, ASC_HIP_ALG_AUTO, staging, {}};
  auto need = asc::HipCsrSpmvAnalysisSize(g.ctx, cv, &opt);
  CHECK(need.ok());
  asc::Buffer analysis = Allocate(g, *need);
  CHECK(asc::HipCsrSpmvAnalyze(g.ctx, cv, Whole(analysis), &opt).ok());
  auto xb = Upload<double>(g, {1, -2, 3, 4, -1});
  auto yb = Upload<double>(g, {10, 20, 30, 40});
  auto x = asc::HipStridedVectorView<const double>::Create(static_cast<const double*>(xb.data()), 5, 1);
  auto y = asc::HipStridedVectorView<double>::Create(static_cast<double*>(yb.data()), 4, 1);
  auto wneed = asc::HipCsrSpmvWorkspaceSize(g.ctx, cv, *x, *y, asc::HipSpmvAlgorithm::kColumnPanel);
  CHECK(wneed.ok());
  asc::Buffer ws = Allocate(g, *wneed + 256);
  auto done = asc::HipCsrSpmv(g.ctx, 2.0, cv, *x, -0.5, *y, Whole(ws),
                              asc::HipSpmvAlgorithm::kColumnPanel);
  CHECK(done.ok() && done->Wait().ok());
  const auto out = Download<double>(g, yb, 4);
  CHECK(out[0] == -9 && out[1] == -26 && out[2] == 7 && out[3] == -24);
  if (staging == ASC_HIP_STAGING_STREAM) {
    // Values changed in place: the stream plan's copy is refreshed, then the
    // plan gives the doubled product (-18, -52, 14, -48 with beta 0).
    std::vector<double> doubled;
    for (double v : kValues4x5) doubled.push_back(2 * v);
    CHECK(Copy(g, const_cast<double*>(cv.values()), MemorySpace::kDevice, doubled.data(),
               MemorySpace::kHost, doubled.size() * sizeof(double)));
    auto refreshed = asc::HipCsrSpmvRefreshValues(g.ctx, cv);
    CHECK(refreshed.ok() && refreshed->Wait().ok());
    done = asc::HipCsrSpmv(g.ctx, 2.0, cv, *x, 0.0, *y, Whole(ws), asc::HipSpmvAlgorithm::kColumnPanel);
    CHECK(done.ok() && done->Wait().ok());
    CHECK((Download<double>(g, yb, 4) == std::vector<double>{-8, -32, 44, -8}));
    // The SpMM's column-panel form on the same analysis reads the current
    // (doubled) values: C = 2A*B with B's rows (1,2,3,4), (0,1,0,-1),
    // (2,0,1,0), (1,1,1,1), (0,0,2,1), every product exact.
    auto bb = Upload<double>(g, {1, 2, 3, 4, 0, 1, 0, -1, 2, 0, 1, 0, 1, 1, 1, 1, 0, 0, 2, 1});
    auto cb = Upload<double>(g, std::vector<double>(16, 0.0));
    const asc::ConstMemoryView bs(bb.data(), 20 * sizeof(double), MemorySpace::kDevice);
    const asc::ConstMemoryView cs(cb.data(), 16 * sizeof(double), MemorySpace::kDevice);
    auto b = asc::SparseBlasMatrixView<const double>::Create(
        static_cast<const double*>(bb.data()), 5, 4, asc::SparseBlasLayout::kRowMajor, 4, bs);
    auto c = asc::SparseBlasMatrixView<double>::Create(static_cast<double*>(cb.data()), 4, 4,
                                                       asc::SparseBlasLayout::kRowMajor, 4, cs);
    CHECK(b.ok() && c.ok());
    auto mneed = asc::HipSpmmWorkspaceSize(g.ctx, asc::SparseBlasTranspose::kNone, cv, *b, *c,
                                           asc::HipSpmmForm::kPanel);
    CHECK(mneed.ok());
    asc::Buffer mws = Allocate(g, *mneed + 256);
    auto spmm = asc::HipSpmm(g.ctx, asc::SparseBlasTranspose::kNone, 1.0, cv, *b, *c, Whole(mws),
                             asc::HipSpmmForm::kPanel);
    CHECK(spmm.ok() && spmm->Wait().ok());
    CHECK((Download<double>(g, cb, 16) ==
           std::vector<double>{2, 6, 10, 14, 0, 8, 0, -8, 26, 28, 68, 66, 0, -4, 24, 16}));
  }
  CHECK(asc_sparse_hip_plan_clear(g.ctx.native()).code == ASC_HIP_OK);
}

// Values-only refresh on trusted device structure (sparse.md:455-466): a
// rebound clone runs every path, keeps its column-panel plan, and a values
// span that runs past its allocation is refused before enqueue.
template <typename E>
void TestRebind(Gpu& g) {
  Case(sizeof(E) == 8 ? "rebound values (double)" : "rebound values (float)");
  auto a = CloneCsr<E>(g, kOffsets4x5, kIndices4x5, Cast<E>(kValues4x5), 4, 5);
  const asc::CsrView<const E> cv = *std::as_const(a).view();
  std::vector<E> doubled;
  for (double v : kValues4x5) doubled.push_back(E(2 * v));
  auto vb = Upload<E>(g, doubled);
  auto rebound = cv.RebindValues(static_cast<const E*>(vb.data()));
  CHECK(rebound.ok() && rebound->canonical_structure_trusted());
  if (!rebound.ok()) return;
  asc_hip_panel_options opt{1, 1, ASC_HIP_ALG_AUTO, ASC_HIP_ALG_AUTO, {}};
  auto need = asc::HipCsrSpmvAnalysisSize(g.ctx, cv, &opt);
  CHECK(need.ok());
  asc::Buffer analysis = Allocate(g, *need);
  CHECK(asc::HipCsrSpmvAnalyze(g.ctx, cv, Whole(analysis), &opt).ok());
  auto xb = Upload<E>(g, {1, -2, 3, 4, -1});
  auto x = asc::HipStridedVectorView<const E>::Create(static_cast<const E*>(xb.data()), 5, 1);
  for (auto alg : {asc::HipSpmvAlgorithm::kOrdered, asc::HipSpmvAlgorithm::kMergePath,
                   asc::HipSpmvAlgorithm::kColumnPanel}) {
    auto yb = Upload<E>(g, {NAN, NAN, NAN, NAN});
    auto y = asc::HipStridedVectorView<E>::Create(static_cast<E*>(yb.data()), 4, 1);
    auto wneed = asc::HipCsrSpmvWorkspaceSize(g.ctx, *rebound, *x, *y, alg);
    CHECK(wneed.ok());
    asc::Buffer ws = Allocate(g, *wneed + 256);
    auto done = asc::HipCsrSpmv(g.ctx, E(1), *rebound, *x, E(0), *y, Whole(ws), alg);
    CHECK(done.ok() && done->Wait().ok());
    CHECK((Download<E>(g, yb, 4) == std::vector<E>{-4, -16, 22, -4}));
  }
  // Values whose eighth element lies one element past the end of a 4 MiB
  // allocation (a size the runtime keeps as requested): the host rebind
  // cannot see the allocation, the operation refuses it and leaves y as it was.
  constexpr std::size_t kBlock = std::size_t{1} << 22;
  asc::Buffer block = Allocate(g, kBlock);
  if (block.data() == nullptr) return;
  const E* tail = reinterpret_cast<const E*>(static_cast<const char*>(block.data()) + kBlock) - 7;
  auto too_short = cv.RebindValues(tail);
  CHECK(too_short.ok());
  auto yb = Upload<E>(g, {1, 2, 3, 4});
  auto y = asc::HipStridedVectorView<E>::Create(static_cast<E*>(yb.data()), 4, 1);
  CHECK_CODE(asc::HipCsrSpmv(g.ctx, E(1), *too_short, *x, E(0), *y, NoWorkspace(),
                             asc::HipSpmvAlgorithm::kOrdered)
                 .status(),
             ErrorCode::kMemoryAccess);
  CHECK((Download<E>(g, yb, 4) == std::vector<E>{1, 2, 3, 4}));
  CHECK(asc_sparse_hip_plan_clear(g.ctx.native()).code == ASC_HIP_OK);
}

// coordinate_test.cc:96-142: triplets (2,1)=5, (0,3)=2, (2,1)=-1, (1,0)=0,
// (0,1)=7 finalize (kSum, kKeep) to (0,1)=7, (0,3)=2, (1,0)=0, (2,1)=4;
// kReject refuses them; (kSum, kDrop) keeps 3 entries.
void TestFinalize(Gpu& g) {
  Case("coordinate finalize golden (double)");
  auto cb = Upload<asc::index_t>(g, {2, 1, 0, 3, 2, 1, 1, 0, 0, 1});
  auto vb = Upload<double>(g, {5, 2, -1, 0, 7});
  const auto* c = static_cast<const asc::index_t*>(cb.data());
  const auto* v = static_cast<const double*>(vb.data());
  using asc::DuplicatePolicy;
  using asc::ExplicitZeroPolicy;
  auto sum = asc::HipFinalizeCoordinates(g.ctx, c, v, 3, 4, 5, DuplicatePolicy::kSum,
                                         ExplicitZeroPolicy::kKeep, *g.device);
  CHECK(sum.ok());
  if (sum.ok()) {
    auto view = sum->view();
    CHECK(view.ok() && view->nnz() == 4);
    std::vector<asc::index_t> pairs(8);
    std::vector<double> vals(4);
    CHECK(FromDevice(g, pairs.data(), view->coordinates(), 8 * 8));
    CHECK(FromDevice(g, vals.data(), view->values(), 4 * 8));
    CHECK((pairs == std::vector<asc::index_t>{0, 1, 0, 3, 1, 0, 2, 1}));
    CHECK((vals == std::vector<double>{7, 2, 0, 4}));
  }
  CHECK_CODE(asc::HipFinalizeCoordinates(g.ctx, c, v, 3, 4, 5, DuplicatePolicy::kReject,
                                         ExplicitZeroPolicy::kKeep, *g.device)
                 .status(),
             ErrorCode::kInvalidArgument);
  auto drop = asc::HipFinalizeCoordinates(g.ctx, c, v, 3, 4, 5, DuplicatePolicy::kSum,
                                          ExplicitZeroPolicy::kDrop, *g.device);
  CHECK(drop.ok() && drop->view()->nnz() == 3);
  // A coordinate outside the shape.
  CHECK_CODE(asc::HipFinalizeCoordinates(g.ctx, c, v, 2, 4, 5, DuplicatePolicy::kSum,
                                         ExplicitZeroPolicy::kKeep, *g.device)
                 .status(),
             ErrorCode::kIndex);
}

// standard_blas_test.cc:69-174 (real parts): indices {0, 2}, values
// {1, -2}, dense {3, 9, 2}.
template <typename E>
void TestLevelOne(Gpu& g) {
  Case(sizeof(E) == 8 ? "Level-1 golden (double)" : "Level-1 golden (float)");
  const std::vector<asc::index_t> idx{0, 2};
  const std::vector<E> vals{1, -2};
  auto host = asc::SparseBlasIndexedVectorView<const E>::Create(
      idx.data(), vals.data(), 2, 3, asc::ConstMemoryView(idx.data(), 16, MemorySpace::kHost),
      asc::ConstMemoryView(vals.data(), 2 * sizeof(E), MemorySpace::kHost));
  CHECK(host.ok());
  auto clone = asc::HipCloneIndexedVector(g.ctx, *host, *g.device);
  CHECK(clone.ok() && clone->completion.Wait().ok());
  auto sp = clone->array.view();
  const asc::SparseBlasIndexedVectorView<const E> csp = *sp;
  auto xb = Upload<E>(g, {3, 9, 2});
  auto rb = Upload<E>(g, {7});
  const asc::ConstMemoryView xs(xb.data(), 3 * sizeof(E), MemorySpace::kDevice);
  const asc::ConstMemoryView rs(rb.data(), sizeof(E), MemorySpace::kDevice);
  auto x = asc::SparseBlasVectorView<E>::Create(static_cast<E*>(xb.data()), 3, 1, xs);
  auto cx = asc::SparseBlasVectorView<const E>::Create(static_cast<const E*>(xb.data()), 3, 1, xs);
  auto r = asc::SparseBlasVectorView<E>::Create(static_cast<E*>(rb.data()), 1, 1, rs);
  for (auto alg : {asc::HipSpmvAlgorithm::kAuto, asc::HipSpmvAlgorithm::kOrdered}) {
    auto d = asc::HipSparseDot(g.ctx, asc::SparseBlasConjugation::kConjugated, csp, *cx, *r, alg);
    CHECK(d.ok() && d->Wait().ok());
    CHECK(Download<E>(g, rb, 1)[0] == E(-1));
  }
  auto a = asc::HipSparseAxpy(g.ctx, E(2), csp, *x);
  CHECK(a.ok() && a->Wait().ok());
  CHECK((Download<E>(g, xb, 3) == std::vector<E>{5, 9, -2}));
  auto gz = asc::HipSparseGatherZero(g.ctx, *x, *sp);
  CHECK(gz.ok() && gz->Wait().ok());
  CHECK((Download<E>(g, xb, 3) == std::vector<E>{0, 9, 0}));
  std::vector<E> gathered(2);
  CHECK(FromDevice(g, gathered.data(), sp->values(), 2 * sizeof(E)));
  CHECK((gathered == std::vector<E>{5, -2}));
  auto sc = asc::HipSparseScatter(g.ctx, csp, *x);
  CHECK(sc.ok() && sc->Wait().ok());
  CHECK((Download<E>(g, xb, 3) == std::vector<E>{5, 9, -2}));
  auto ga = asc::HipSparseGather(g.ctx, *cx, *sp);
  CHECK(ga.ok() && ga->Wait().ok());
  // An untrusted device view is refused (blas.h:1231-1239 for the host form).
  auto untrusted = asc::SparseBlasIndexedVectorView<const E>::Create(
      sp->indices(), sp->values(), 2, 3, asc::ConstMemoryView(sp->indices(), 16, MemorySpace::kDevice),
      asc::ConstMemoryView(sp->values(), 2 * sizeof(E), MemorySpace::kDevice));
  CHECK(untrusted.ok());
  CHECK_CODE(asc::HipSparseAxpy(g.ctx, E(1), *untrusted, *x).status(), ErrorCode::kInvalidArgument);
}

// standard_blas_test.cc:335-398 (real parts): A = [[2, 0], [1, 3]] lower.
template <typename E>
void TestTriangular(Gpu& g) {
  Case(sizeof(E) == 8 ? "triangular solve golden (double)" : "triangular solve golden (float)");
  auto a = CloneCsr<E>(g, {0, 1, 3}, {0, 0, 1}, {2, 1, 3}, 2, 2);
  const asc::CsrView<const E> cv = *std::as_const(a).view();
  auto plan = asc::HipTriangularPlan<E, asc::SparseCompressedFormat::kCsr>::Analyze(
      g.ctx, cv, asc::SparseBlasTriangle::kLower, asc::SparseBlasDiagonal::kNonUnit,
      asc::SparseBlasTranspose::kNone, *g.device);
  CHECK(plan.ok());
  if (!plan.ok()) return;
  CHECK(plan->levels() == 2);
  auto xb = Upload<E>(g, {4, -1});
  const asc::ConstMemoryView xs(xb.data(), 2 * sizeof(E), MemorySpace::kDevice);
  auto x = asc::SparseBlasVectorView<E>::Create(static_cast<E*>(xb.data()), 2, 1, xs);
  auto done = plan->Solve(g.ctx, E(1), *x);
  CHECK(done.ok() && done->Wait().ok());
  CHECK((Download<E>(g, xb, 2) == std::vector<E>{2, -1}));
  // Several right-hand sides, row-major: [[4, 2], [-1, 10]] -> [[2, 1], [-1, 3]].
  auto bb = Upload<E>(g, {4, 2, -1, 10});
  const asc::ConstMemoryView bs(bb.data(), 4 * sizeof(E), MemorySpace::kDevice);
  auto bm = asc::SparseBlasMatrixView<E>::Create(static_cast<E*>(bb.data()), 2, 2,
                                                 asc::SparseBlasLayout::kRowMajor, 2, bs);
  done = plan->Solve(g.ctx, E(1), *bm);
  CHECK(done.ok() && done->Wait().ok());
  CHECK((Download<E>(g, bb, 4) == std::vector<E>{2, 1, -1, 3}));
  // A singular non-unit diagonal is refused.
  auto s = CloneCsr<E>(g, {0, 1, 2}, {0, 1}, {1, 0}, 2, 2);
  const asc::CsrView<const E> sv = *std::as_const(s).view();
  CHECK_CODE((asc::HipTriangularPlan<E, asc::SparseCompressedFormat::kCsr>::Analyze(
                  g.ctx, sv, asc::SparseBlasTriangle::kLower, asc::SparseBlasDiagonal::kNonUnit,
                  asc::SparseBlasTranspose::kNone, *g.device)
                  .status()),
             ErrorCode::kInvalidArgument);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "validate";
  TestHostViewValidation();
  TestRejectionMatrix();
  if (mode == "gpu") {
    auto count = asc::HipDeviceCount();
    if (!count.ok() || *count < 1) {
      std::fprintf(stderr, "FAILED: no HIP device\n");
      return 2;
    }
    auto exec = asc::CreateHipExecutionContext(0);
    CHECK(exec.ok() && exec->backend() == asc::Backend::kHip);
    auto ctx = asc::SparseHipContext::Create(*exec);
    CHECK(ctx.ok());
    if (!ctx.ok()) return 3;
    auto res = asc::HipMemoryResource::Create(0, MemorySpace::kDevice);
    CHECK(res.ok());
    if (!res.ok()) return 3;
    Gpu g{std::move(*ctx), std::move(*res)};
    TestCoreProvider(g);
    TestCsrSpmvGolden<double>(g);
    TestCsrSpmvGolden<float>(g);
    TestSummationOrder(g);
    TestStandard<double>(g);
    TestStandard<float>(g);
    TestConversions<double>(g);
    TestConversions<float>(g);
    TestColumnPanel(g, ASC_HIP_STAGING_AUTO);
    TestColumnPanel(g, ASC_HIP_STAGING_STREAM);
    TestFinalize(g);
    TestLevelOne<double>(g);
    TestLevelOne<float>(g);
    TestTriangular<double>(g);
    TestTriangular<float>(g);
    TestRebind<double>(g);
    TestRebind<float>(g);
  }
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
