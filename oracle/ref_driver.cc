// C entry points over the reference's own serial Sparse BLAS and conversions,
// compiled from the reference sources by oracle/build_ref.sh into
// oracle/_ref/libref_driver.so.  TEST INFRASTRUCTURE ONLY: it pins the C
// oracle (oracle.c) against the real reference and can serve as the
// benchmark's CPU baseline; nothing in the product loads it.
//
// The reference's compat Spmv reads x and writes y through the expression
// adapter protocol (include/asc/expression/expression.h:118-224,
// writable.h:138-304); RefStrided below is a minimal strided host vector
// that implements that protocol.
#include <array>
#include <cstdint>
#include <cstring>
#include <span>
#include <type_traits>

#include "asc/core/execution.h"
#include "asc/core/extents.h"
#include "asc/core/memory.h"
#include "asc/core/status.h"
#include "asc/core/types.h"
#include "asc/expression/expression.h"
#include "asc/expression/writable.h"
#include "asc/sparse/blas.h"
#include "asc/sparse/compressed.h"
#include "asc/sparse/coordinate.h"

template <typename E>
struct RefStrided {
  E* data = nullptr;
  asc::extent_t size = 0;
  asc::stride_t stride = 1;
};

namespace asc {

template <typename E>
struct ExpressionAdapter<::RefStrided<E>> {
  using value_type = std::remove_const_t<E>;
  static constexpr rank_t rank = 1;
  static constexpr SparsityEffect sparsity_effect = SparsityEffect::kStructurePreserving;
  static constexpr ExpressionOperation operation = ExpressionOperation::kExternal;
  static constexpr std::array<extent_t, 1> Shape(const ::RefStrided<E>& v) noexcept {
    return {v.size};
  }
  static constexpr value_type Read(const ::RefStrided<E>& v,
                                   std::span<const index_t, 1> c) noexcept {
    return v.data[c[0] * v.stride];
  }
  static bool MayAlias(const ::RefStrided<E>& v, AliasToken token) noexcept {
    if (v.data == nullptr || v.size <= 0 || token.identity() == nullptr) return false;
    const auto b = reinterpret_cast<std::uintptr_t>(v.data);
    const auto n = static_cast<std::size_t>((v.size - 1) * v.stride + 1) * sizeof(value_type);
    const auto a = reinterpret_cast<std::uintptr_t>(token.identity());
    return a >= b && a < b + n;
  }
  static Status ValidateAccess(const ::RefStrided<E>&, const ExecutionContext& context) {
    return context.backend() == Backend::kSerial
               ? Status::Ok()
               : Status(ErrorCode::kUnsupported, "serial only");
  }
};

template <typename E>
struct ExpressionPlacementAdapter<::RefStrided<E>> {
  static constexpr MemorySpace Space(const ::RefStrided<E>&) noexcept {
    return MemorySpace::kHost;
  }
  static constexpr ExpressionAliasMetadata Alias(const ::RefStrided<E>& v) noexcept {
    const std::size_t n =
        v.size <= 0 ? 0
                    : static_cast<std::size_t>((v.size - 1) * v.stride + 1) *
                          sizeof(std::remove_const_t<E>);
    return ExpressionAliasMetadata(v.data, v.data, n);
  }
};

template <typename E>
  requires(!std::is_const_v<E>)
struct WritableExpressionAdapter<::RefStrided<E>> {
  using value_type = std::remove_const_t<E>;
  static constexpr std::array<extent_t, 1> Shape(const ::RefStrided<E>& v) noexcept {
    return {v.size};
  }
  static constexpr ExpressionAliasMetadata Alias(const ::RefStrided<E>& v) noexcept {
    return ExpressionPlacementAdapter<::RefStrided<E>>::Alias(v);
  }
  static constexpr bool IsUnique(const ::RefStrided<E>& v) noexcept {
    return v.size <= 1 || v.stride != 0;
  }
  static constexpr void Write(::RefStrided<E>& v, std::span<const index_t, 1> c,
                              value_type value) noexcept {
    v.data[c[0] * v.stride] = value;
  }
  static Status ValidateAccess(const ::RefStrided<E>& v, const ExecutionContext& context) {
    return ExpressionAdapter<::RefStrided<E>>::ValidateAccess(v, context);
  }
};

}  // namespace asc

namespace {

template <typename T>
int compat_spmv(int64_t rows, int64_t cols, int64_t nnz, const int64_t* off,
                const int64_t* idx, const T* val, T alpha, const T* x, int64_t incx,
                T beta, T* y, int64_t incy) {
  const std::array<asc::extent_t, 2> shape{rows, cols};
  auto view = asc::CsrView<const T>::Create(off, idx, val, shape, nnz, asc::MemorySpace::kHost);
  if (!view.ok()) return static_cast<int>(view.status().code());
  const RefStrided<const T> in{x, cols, incx};
  RefStrided<T> out{y, rows, incy};
  const asc::Status st =
      asc::Spmv(asc::ExecutionContext::Serial(), alpha, *view, in, beta, out);
  return static_cast<int>(st.code());
}

template <typename T, asc::SparseCompressedFormat F>
int std_spmv_typed(int op, int64_t rows, int64_t cols, int64_t nnz, const int64_t* off,
                   const int64_t* idx, const T* val, T alpha, const T* x, int64_t nx,
                   int64_t incx, T* y, int64_t ny, int64_t incy) {
  const std::array<asc::extent_t, 2> shape{rows, cols};
  auto view = asc::CompressedSparseView<const T, F>::Create(off, idx, val, shape, nnz,
                                                            asc::MemorySpace::kHost);
  if (!view.ok()) return static_cast<int>(view.status().code());
  const std::size_t xb = nx > 0 ? static_cast<std::size_t>((nx - 1) * (incx < 0 ? -incx : incx) + 1) * sizeof(T) : 0;
  const std::size_t yb = ny > 0 ? static_cast<std::size_t>((ny - 1) * (incy < 0 ? -incy : incy) + 1) * sizeof(T) : 0;
  const T* xlo = incx < 0 && nx > 0 ? x + (nx - 1) * incx : x;
  T* ylo = incy < 0 && ny > 0 ? y + (ny - 1) * incy : y;
  auto in = asc::SparseBlasVectorView<const T>::Create(
      x, nx, incx, asc::ConstMemoryView(xlo, xb, asc::MemorySpace::kHost));
  auto out = asc::SparseBlasVectorView<T>::Create(
      y, ny, incy, asc::ConstMemoryView(ylo, yb, asc::MemorySpace::kHost));
  if (!in.ok()) return static_cast<int>(in.status().code());
  if (!out.ok()) return static_cast<int>(out.status().code());
  return static_cast<int>(asc::Spmv(asc::ExecutionContext::Serial(),
                                    static_cast<asc::SparseBlasTranspose>(op), alpha, *view,
                                    *in, *out)
                              .code());
}

template <typename T, asc::SparseCompressedFormat F>
int std_spmm_typed(int op, int64_t rows, int64_t cols, int64_t nnz, const int64_t* off,
                   const int64_t* idx, const T* val, T alpha, const T* b, int b_layout,
                   int64_t b_rows, int64_t ldb, int64_t k, T* c, int c_layout,
                   int64_t c_rows, int64_t ldc) {
  const std::array<asc::extent_t, 2> shape{rows, cols};
  auto view = asc::CompressedSparseView<const T, F>::Create(off, idx, val, shape, nnz,
                                                            asc::MemorySpace::kHost);
  if (!view.ok()) return static_cast<int>(view.status().code());
  auto bytes = [](int layout, int64_t r, int64_t cc, int64_t ld) -> std::size_t {
    if (r == 0 || cc == 0) return 0;
    const int64_t major = layout == 1 ? r : cc, minor = layout == 1 ? cc : r;
    return static_cast<std::size_t>((major - 1) * ld + minor) * sizeof(T);
  };
  auto in = asc::SparseBlasMatrixView<const T>::Create(
      b, b_rows, k, static_cast<asc::SparseBlasLayout>(b_layout), ldb,
      asc::ConstMemoryView(b, bytes(b_layout, b_rows, k, ldb), asc::MemorySpace::kHost));
  auto out = asc::SparseBlasMatrixView<T>::Create(
      c, c_rows, k, static_cast<asc::SparseBlasLayout>(c_layout), ldc,
      asc::ConstMemoryView(c, bytes(c_layout, c_rows, k, ldc), asc::MemorySpace::kHost));
  if (!in.ok()) return static_cast<int>(in.status().code());
  if (!out.ok()) return static_cast<int>(out.status().code());
  return static_cast<int>(asc::Spmm(asc::ExecutionContext::Serial(),
                                    static_cast<asc::SparseBlasTranspose>(op), alpha, *view,
                                    *in, *out)
                              .code());
}

using Dyn2 = asc::Extents<asc::kDynamicExtent, asc::kDynamicExtent>;

template <typename Owner>
int copy_out(Owner& owner, int64_t outer, int64_t* off, int64_t* idx, double* val) {
  auto view = owner.view();
  if (!view.ok()) return static_cast<int>(view.status().code());
  std::memcpy(off, view->outer_offsets(), sizeof(int64_t) * (outer + 1));
  std::memcpy(idx, view->inner_indices(), sizeof(int64_t) * view->nnz());
  std::memcpy(val, view->values(), sizeof(double) * view->nnz());
  return 0;
}

}  // namespace

extern "C" {

int ref_compat_spmv_f64(int64_t rows, int64_t cols, int64_t nnz, const int64_t* off,
                        const int64_t* idx, const double* val, double alpha,
                        const double* x, int64_t incx, double beta, double* y,
                        int64_t incy) {
  return compat_spmv<double>(rows, cols, nnz, off, idx, val, alpha, x, incx, beta, y, incy);
}

int ref_compat_spmv_f32(int64_t rows, int64_t cols, int64_t nnz, const int64_t* off,
                        const int64_t* idx, const float* val, float alpha, const float* x,
                        int64_t incx, float beta, float* y, int64_t incy) {
  return compat_spmv<float>(rows, cols, nnz, off, idx, val, alpha, x, incx, beta, y, incy);
}

int ref_std_spmv_f64(int fmt, int op, int64_t rows, int64_t cols, int64_t nnz,
                     const int64_t* off, const int64_t* idx, const double* val,
                     double alpha, const double* x, int64_t incx, double* y, int64_t incy) {
  const int64_t nx = op == 0 ? cols : rows, ny = op == 0 ? rows : cols;
  if (fmt == 0)
    return std_spmv_typed<double, asc::SparseCompressedFormat::kCsr>(
        op, rows, cols, nnz, off, idx, val, alpha, x, nx, incx, y, ny, incy);
  return std_spmv_typed<double, asc::SparseCompressedFormat::kCsc>(
      op, rows, cols, nnz, off, idx, val, alpha, x, nx, incx, y, ny, incy);
}

int ref_std_spmm_f64(int fmt, int op, int64_t rows, int64_t cols, int64_t nnz,
                     const int64_t* off, const int64_t* idx, const double* val,
                     double alpha, const double* b, int b_layout, int64_t ldb, int64_t k,
                     double* c, int c_layout, int64_t ldc) {
  const int64_t br = op == 0 ? cols : rows, cr = op == 0 ? rows : cols;
  if (fmt == 0)
    return std_spmm_typed<double, asc::SparseCompressedFormat::kCsr>(
        op, rows, cols, nnz, off, idx, val, alpha, b, b_layout, br, ldb, k, c, c_layout, cr, ldc);
  return std_spmm_typed<double, asc::SparseCompressedFormat::kCsc>(
      op, rows, cols, nnz, off, idx, val, alpha, b, b_layout, br, ldb, k, c, c_layout, cr, ldc);
}

int ref_coo_to_compressed_f64(int fmt, int64_t rows, int64_t cols, int64_t nnz,
                              const int64_t* coords, const double* values, int64_t* off,
                              int64_t* idx, double* val) {
  asc::HostMemoryResource resource;
  auto extents = Dyn2::Create(rows, cols);
  if (!extents.ok()) return static_cast<int>(extents.status().code());
  auto coo = asc::CoordinateArray<double, Dyn2>::Create(
      resource, *extents, std::span<const asc::index_t>(coords, 2 * nnz),
      std::span<const double>(values, nnz));
  if (!coo.ok()) return static_cast<int>(coo.status().code());
  auto view = coo->view();
  if (!view.ok()) return static_cast<int>(view.status().code());
  const asc::ExecutionContext ctx = asc::ExecutionContext::Serial();
  if (fmt == 0) {
    auto out = asc::ConvertToCsr<double>(ctx, *view, resource);
    if (!out.ok()) return static_cast<int>(out.status().code());
    return copy_out(*out, rows, off, idx, val);
  }
  auto out = asc::ConvertToCsc<double>(ctx, *view, resource);
  if (!out.ok()) return static_cast<int>(out.status().code());
  return copy_out(*out, cols, off, idx, val);
}

int ref_transpose_f64(int src_fmt, int64_t rows, int64_t cols, const int64_t* off,
                      const int64_t* idx, const double* values, int64_t* out_off,
                      int64_t* out_idx, double* out_val) {
  asc::HostMemoryResource resource;
  const asc::ExecutionContext ctx = asc::ExecutionContext::Serial();
  const int64_t src_outer = src_fmt == 0 ? rows : cols;
  const int64_t nnz = off[src_outer];
  const std::span<const asc::nnz_t> so(off, src_outer + 1);
  const std::span<const asc::index_t> si(idx, nnz);
  const std::span<const double> sv(values, nnz);
  if (src_fmt == 0) {
    auto csr = asc::CsrArray<double>::Create(resource, rows, cols, so, si, sv);
    if (!csr.ok()) return static_cast<int>(csr.status().code());
    auto view = csr->view();
    if (!view.ok()) return static_cast<int>(view.status().code());
    auto out = asc::ConvertToCsc<double>(ctx, *view, resource);
    if (!out.ok()) return static_cast<int>(out.status().code());
    return copy_out(*out, cols, out_off, out_idx, out_val);
  }
  auto csc = asc::CscArray<double>::Create(resource, rows, cols, so, si, sv);
  if (!csc.ok()) return static_cast<int>(csc.status().code());
  auto view = csc->view();
  if (!view.ok()) return static_cast<int>(view.status().code());
  auto out = asc::ConvertToCsr<double>(ctx, *view, resource);
  if (!out.ok()) return static_cast<int>(out.status().code());
  return copy_out(*out, rows, out_off, out_idx, out_val);
}

}  // extern "C"

namespace {

template <typename T>
int RefFinalize(int64_t rows, int64_t cols, int64_t nnz, const int64_t* coords,
                const T* values, int duplicate_policy, int zero_policy,
                int64_t* out_coords, T* out_values, int64_t* out_nnz) {
  asc::HostMemoryResource resource;
  auto extents = Dyn2::Create(rows, cols);
  if (!extents.ok()) return static_cast<int>(extents.status().code());
  auto builder = asc::CoordinateBuilder<T, Dyn2>::Create(resource, *extents, nnz);
  if (!builder.ok()) return static_cast<int>(builder.status().code());
  for (int64_t p = 0; p < nnz; ++p) {
    const std::array<asc::index_t, 2> c{coords[2 * p], coords[2 * p + 1]};
    const asc::Status st = builder->Add(std::span<const asc::index_t, 2>(c), values[p]);
    if (!st.ok()) return static_cast<int>(st.code());
  }
  auto fin = std::move(*builder).Finalize(asc::ExecutionContext::Serial(),
                                          static_cast<asc::DuplicatePolicy>(duplicate_policy),
                                          static_cast<asc::ExplicitZeroPolicy>(zero_policy));
  if (!fin.ok()) return static_cast<int>(fin.status().code());
  auto view = fin->view();
  if (!view.ok()) return static_cast<int>(view.status().code());
  *out_nnz = view->nnz();
  std::memcpy(out_coords, view->coordinates(), sizeof(int64_t) * 2 * view->nnz());
  std::memcpy(out_values, view->values(), sizeof(T) * view->nnz());
  return 0;
}

}  // namespace

extern "C" {

int ref_finalize_f64(int64_t rows, int64_t cols, int64_t nnz, const int64_t* coords,
                     const double* values, int duplicate_policy, int zero_policy,
                     int64_t* out_coords, double* out_values, int64_t* out_nnz) {
  return RefFinalize<double>(rows, cols, nnz, coords, values, duplicate_policy, zero_policy,
                             out_coords, out_values, out_nnz);
}

int ref_finalize_f32(int64_t rows, int64_t cols, int64_t nnz, const int64_t* coords,
                     const float* values, int duplicate_policy, int zero_policy,
                     int64_t* out_coords, float* out_values, int64_t* out_nnz) {
  return RefFinalize<float>(rows, cols, nnz, coords, values, duplicate_policy, zero_policy,
                            out_coords, out_values, out_nnz);
}

}  // extern "C"

// ---- sparse Level-1: the reference's own SparseDot / SparseAxpy /
// SparseGather / SparseGatherZero / SparseScatter (blas.h:1330-1582) on host
// arrays (dense stride `inc`).  op: 0 dot, 1 axpy, 2 gather, 3 gather-zero,
// 4 scatter.  `v` is read (dot, axpy, scatter) or written (gathers); `x` is
// read (dot, gather) or written (axpy, gather-zero, scatter); the dot result
// goes to *dot.
namespace {

template <typename T>
int RefLevel1(int op, int64_t nnz, int64_t extent, const int64_t* idx, T* v, T* x, int64_t inc,
              T alpha, T* dot) {
  const asc::ExecutionContext ctx = asc::ExecutionContext::Serial();
  const asc::ConstMemoryView istore{idx, sizeof(int64_t) * static_cast<std::size_t>(nnz),
                                    asc::MemorySpace::kHost};
  const asc::ConstMemoryView vstore{v, sizeof(T) * static_cast<std::size_t>(nnz),
                                    asc::MemorySpace::kHost};
  const std::size_t xbytes =
      extent > 0 ? sizeof(T) * static_cast<std::size_t>((extent - 1) * inc + 1) : 0;
  const asc::ConstMemoryView xstore{x, xbytes, asc::MemorySpace::kHost};
  auto sparse = asc::SparseBlasIndexedVectorView<T>::Create(idx, v, nnz, extent, istore, vstore);
  auto dense = asc::SparseBlasVectorView<T>::Create(x, extent, inc, xstore);
  if (!sparse.ok()) return static_cast<int>(sparse.status().code());
  if (!dense.ok()) return static_cast<int>(dense.status().code());
  const asc::SparseBlasIndexedVectorView<const T> csparse(*sparse);
  const asc::SparseBlasVectorView<const T> cdense(*dense);
  asc::Status st;
  switch (op) {
    case 0: {
      auto r = asc::SparseDot(ctx, asc::SparseBlasConjugation::kUnconjugated, csparse, cdense);
      if (!r.ok()) return static_cast<int>(r.status().code());
      *dot = *r;
      return 0;
    }
    case 1: st = asc::SparseAxpy(ctx, alpha, csparse, *dense); break;
    case 2: st = asc::SparseGather(ctx, cdense, *sparse); break;
    case 3: st = asc::SparseGatherZero(ctx, *dense, *sparse); break;
    default: st = asc::SparseScatter(ctx, csparse, *dense); break;
  }
  return static_cast<int>(st.code());
}

}  // namespace

extern "C" {

int ref_level1_f64(int op, int64_t nnz, int64_t extent, const int64_t* idx, double* v, double* x,
                   int64_t inc, double alpha, double* dot) {
  return RefLevel1<double>(op, nnz, extent, idx, v, x, inc, alpha, dot);
}

int ref_level1_f32(int op, int64_t nnz, int64_t extent, const int64_t* idx, float* v, float* x,
                   int64_t inc, float alpha, float* dot) {
  return RefLevel1<float>(op, nnz, extent, idx, v, x, inc, alpha, dot);
}

}  // extern "C"

// ---- triangular solve: the reference's SparseTriangularSolveMultiple
// (blas.h:1823-1880) on host arrays; b is n x k, row-major (ld = k) when
// row_major, else column-major (ld = n).
namespace {

template <typename T, asc::SparseCompressedFormat F>
int RefTrsv(int op, int lower, int unit, int64_t n, int64_t nnz, const int64_t* off,
            const int64_t* idx, const T* val, T alpha, T* b, int64_t k, int row_major) {
  const std::array<asc::extent_t, 2> shape{n, n};
  auto m = asc::CompressedSparseView<const T, F>::Create(off, idx, val, shape, nnz,
                                                         asc::MemorySpace::kHost);
  if (!m.ok()) return static_cast<int>(m.status().code());
  auto tri = asc::SparseBlasTriangularView<T, F>::Create(
      *m, lower ? asc::SparseBlasTriangle::kLower : asc::SparseBlasTriangle::kUpper,
      unit ? asc::SparseBlasDiagonal::kUnit : asc::SparseBlasDiagonal::kNonUnit);
  if (!tri.ok()) return static_cast<int>(tri.status().code());
  const asc::MutableMemoryView storage{b, sizeof(T) * static_cast<std::size_t>(n * k),
                                       asc::MemorySpace::kHost};
  auto rhs = asc::SparseBlasMatrixView<T>::Create(
      b, n, k, row_major ? asc::SparseBlasLayout::kRowMajor : asc::SparseBlasLayout::kColumnMajor,
      row_major ? (k > 0 ? k : 1) : (n > 0 ? n : 1), storage);
  if (!rhs.ok()) return static_cast<int>(rhs.status().code());
  const auto t = op == 0 ? asc::SparseBlasTranspose::kNone
                         : (op == 1 ? asc::SparseBlasTranspose::kTranspose
                                    : asc::SparseBlasTranspose::kConjugateTranspose);
  return static_cast<int>(
      asc::SparseTriangularSolveMultiple(asc::ExecutionContext::Serial(), t, alpha, *tri, *rhs)
          .code());
}

}  // namespace

extern "C" {

int ref_trsv_f64(int fmt, int op, int lower, int unit, int64_t n, int64_t nnz, const int64_t* off,
                 const int64_t* idx, const double* val, double alpha, double* b, int64_t k,
                 int row_major) {
  if (fmt == 0)
    return RefTrsv<double, asc::SparseCompressedFormat::kCsr>(op, lower, unit, n, nnz, off, idx,
                                                              val, alpha, b, k, row_major);
  return RefTrsv<double, asc::SparseCompressedFormat::kCsc>(op, lower, unit, n, nnz, off, idx, val,
                                                            alpha, b, k, row_major);
}

int ref_trsv_f32(int fmt, int op, int lower, int unit, int64_t n, int64_t nnz, const int64_t* off,
                 const int64_t* idx, const float* val, float alpha, float* b, int64_t k,
                 int row_major) {
  if (fmt == 0)
    return RefTrsv<float, asc::SparseCompressedFormat::kCsr>(op, lower, unit, n, nnz, off, idx, val,
                                                             alpha, b, k, row_major);
  return RefTrsv<float, asc::SparseCompressedFormat::kCsc>(op, lower, unit, n, nnz, off, idx, val,
                                                           alpha, b, k, row_major);
}

}  // extern "C"
