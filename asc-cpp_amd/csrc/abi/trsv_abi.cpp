// C ABI: level-scheduled sparse triangular solve (blas.h:1760-1880),
// replacing StandardTriangularSolveErased (providers/cuda.h:250-253).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <string>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using namespace asc_hip_abi;

namespace {

constexpr int64_t kU32Max = std::numeric_limits<uint32_t>::max();
// Triangular schedules a context keeps (the oldest goes first).
constexpr size_t kMaxSchedules = 64;

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// op(A)'s rows are the stored outer segments for CSR with op N and CSC with
// op T/H; the other two forms transpose the storage into the analysis.
bool needs_copy(const asc_hip_sparse_desc& a, uint8_t transpose) {
  const bool csr = a.format == ASC_HIP_CSR;
  return (csr && transpose != ASC_HIP_OP_N) || (!csr && transpose == ASC_HIP_OP_N);
}

struct TrsvLayout {
  size_t level_ptr = 0, row_order = 0, copy_offsets = 0, copy_indices = 0, copy_values = 0;
  size_t scratch = 0, total = 0;
  bool copy = false;
};

TrsvLayout trsv_layout(const asc_hip_sparse_desc& a, uint8_t transpose) {
  TrsvLayout L;
  const size_t n = static_cast<size_t>(a.rows), nnz = static_cast<size_t>(a.nnz);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  L.level_ptr = take(sizeof(int64_t) * (n + 1));
  L.row_order = take(sizeof(uint32_t) * (n > 0 ? n : 1));
  L.copy = needs_copy(a, transpose);
  if (L.copy) {
    L.copy_offsets = take(sizeof(int64_t) * (n + 1));
    L.copy_indices = take(static_cast<size_t>(a.index_bytes) * (nnz > 0 ? nnz : 1));
    L.copy_values = take(elem_size(a.element_kind) * (nnz > 0 ? nnz : 1));
  }
  size_t scratch = asc_hip::trsv_analysis_scratch_bytes(a.rows);
  if (L.copy) scratch = std::max(scratch, asc_hip::transpose_workspace_bytes(as_compressed(a, false)));
  L.scratch = take(scratch);
  L.total = off;
  return L;
}

asc_hip_status check_trsv_matrix(asc_hip_context* ctx, uint8_t transpose,
                                 const asc_hip_sparse_desc* a) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (transpose > ASC_HIP_OP_H) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown transpose");
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "triangular solve requires CSR or CSC");
  if (a->element_kind != ASC_HIP_F32 && a->element_kind != ASC_HIP_F64)
    return fail(ASC_HIP_UNSUPPORTED, "triangular solve supports real FP32/FP64 storage");
  if (a->rows != a->cols) return fail(ASC_HIP_SHAPE, "A triangular sparse matrix must be square");
  if (a->rows > kU32Max - 1 || a->nnz > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "device triangular solve supports extents and nnz below 2^32");
  return ok();
}

}  // namespace

extern "C" {

asc_hip_status asc_sparse_hip_triangular_analysis_size(asc_hip_context* ctx, uint8_t transpose,
                                                       const asc_hip_sparse_desc* a,
                                                       size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  asc_hip_status st = check_trsv_matrix(ctx, transpose, a);
  if (st.code) return st;
  *out = trsv_layout(*a, transpose).total;
  return ok();
}

asc_hip_status asc_sparse_hip_triangular_analyze(asc_hip_context* ctx, uint8_t transpose,
                                                 uint8_t triangle, uint8_t diagonal,
                                                 const asc_hip_sparse_desc* a, void* analysis,
                                                 size_t analysis_bytes,
                                                 asc_hip_triangular_info* info) {
  asc_hip_status st = check_trsv_matrix(ctx, transpose, a);
  if (st.code) return st;
  if (triangle > ASC_HIP_LOWER) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown triangle");
  if (diagonal > ASC_HIP_UNIT) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown diagonal");
  if (info == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output info is null");
  const TrsvLayout L = trsv_layout(*a, transpose);
  const SparseSpans sp = sparse_spans(*a);
  if ((st = check_workspace(ctx, analysis, analysis_bytes, L.total, {sp.first, sp.second, sp.values})).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  char* w = static_cast<char*>(analysis);
  const asc_hip::Compressed stored = as_compressed(*a, false);
  asc_hip::Compressed rows = stored;
  if (L.copy) {
    rows.offsets = reinterpret_cast<const int64_t*>(w + L.copy_offsets);
    rows.indices = w + L.copy_indices;
    rows.values = w + L.copy_values;
    rows.outer = stored.inner;
    rows.inner = stored.outer;
    const hipError_t e = asc_hip::transpose_launch(
        stored, reinterpret_cast<int64_t*>(w + L.copy_offsets), w + L.copy_indices,
        a->index_bytes, w + L.copy_values, w + L.scratch, ctx->stream);
    if ((st = finish_launch(ctx, e, "triangular transpose")).code) return st;
  }
  // The stored triangle as an index relation: CSR keeps (row, col) as
  // (outer, inner), CSC as (inner, outer).
  const bool lower = triangle == ASC_HIP_LOWER;
  const bool inner_le_outer = (a->format == ASC_HIP_CSR) == lower;
  const bool solve_lower = transpose == ASC_HIP_OP_N ? lower : !lower;  // blas.h:1794-1796
  asc_hip::TrsvLevels lv;
  int failure = 0;
  const hipError_t e = asc_hip::trsv_analysis_launch(
      stored, inner_le_outer, diagonal == ASC_HIP_UNIT, rows, solve_lower,
      reinterpret_cast<int64_t*>(w + L.level_ptr), reinterpret_cast<uint32_t*>(w + L.row_order),
      w + L.scratch, &lv, &failure, ctx->stream);
  if ((st = finish_launch(ctx, e, "triangular analysis")).code) return st;
  if (failure == 1)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "Sparse triangular storage contains an entry outside the declared triangle");
  if (failure == 2) return fail(ASC_HIP_INVALID_ARGUMENT, "A non-unit sparse triangular matrix is singular");
  *info = asc_hip_triangular_info{};
  info->rows = a->rows;
  info->nnz = a->nnz;
  info->levels = lv.levels;
  info->max_level_rows = lv.max_level_rows;
  info->offsets = a->outer_offsets;
  info->indices = a->inner_indices;
  info->values = a->values;
  info->transpose = transpose;
  info->triangle = triangle;
  info->diagonal = diagonal;
  info->format = a->format;
  {
    std::lock_guard<std::mutex> lock(ctx->plan_mutex);
    auto& list = ctx->trsv_schedules;
    for (size_t i = 0; i < list.size(); ++i)
      if (list[i].analysis == analysis) {
        list.erase(list.begin() + static_cast<std::ptrdiff_t>(i));
        break;
      }
    if (list.size() >= kMaxSchedules) list.erase(list.begin());  // oldest first
    list.push_back({analysis, lv.levels, std::move(lv.host_level_ptr)});
  }
  return ok();
}

asc_hip_status asc_sparse_hip_triangular_solve(asc_hip_context* ctx, uint8_t transpose,
                                               asc_hip_scalar alpha,
                                               const asc_hip_sparse_desc* a,
                                               const asc_hip_matrix_desc* rhs,
                                               const void* analysis, size_t analysis_bytes,
                                               const asc_hip_triangular_info* info) {
  asc_hip_status st = check_trsv_matrix(ctx, transpose, a);
  if (st.code) return st;
  if (info == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "triangular info is null");
  if (info->rows != a->rows || info->nnz != a->nnz || info->offsets != a->outer_offsets ||
      info->indices != a->inner_indices || info->values != a->values ||
      info->transpose != transpose || info->format != a->format)
    return fail(ASC_HIP_INVALID_ARGUMENT, "the triangular analysis belongs to another matrix or operation");
  if ((st = check_matrix(ctx, rhs, a->element_kind, a->rows, -1, true, "right-hand sides")).code)
    return st;
  const Span rs = matrix_span(*rhs);
  if (sparse_overlaps(*a, rs))
    return fail(ASC_HIP_INVALID_ARGUMENT, "Sparse triangular right-hand side overlaps the matrix");
  const TrsvLayout L = trsv_layout(*a, transpose);
  if (analysis == nullptr || analysis_bytes < L.total)
    return fail(ASC_HIP_INVALID_ARGUMENT, "analysis buffer is smaller than the analysis size");
  if ((st = check_device_pointer(ctx, analysis, "analysis", L.total)).code) return st;
  if (overlaps(span_of(analysis, L.total), rs))
    return fail(ASC_HIP_INVALID_ARGUMENT, "right-hand sides overlap the analysis");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  const bool row_major = rhs->layout == ASC_HIP_ROW_MAJOR;
  const int64_t rstride = row_major ? rhs->leading_dimension : 1;
  const int64_t cstride = row_major ? 1 : rhs->leading_dimension;
  void* x = const_cast<void*>(rhs->data);
  const asc_hip::Elem elem = static_cast<asc_hip::Elem>(a->element_kind);
  if (alpha.real == 0.0)  // blas.h:1787-1792: zero the right-hand sides
    return finish_launch(ctx, asc_hip::dense_zero_launch(elem, x, rhs->rows, rhs->cols, rstride,
                                                         cstride, ctx->stream),
                         "triangular zero");
  const char* w = static_cast<const char*>(analysis);
  asc_hip::Compressed rows = as_compressed(*a, false);
  if (L.copy) {
    rows.offsets = reinterpret_cast<const int64_t*>(w + L.copy_offsets);
    rows.indices = w + L.copy_indices;
    rows.values = w + L.copy_values;
    std::swap(rows.outer, rows.inner);
  }
  const bool lower = info->triangle == ASC_HIP_LOWER;
  const bool solve_lower = transpose == ASC_HIP_OP_N ? lower : !lower;
  asc_hip::TrsvLevels lv;
  lv.levels = info->levels;
  lv.max_level_rows = info->max_level_rows;
  {
    std::lock_guard<std::mutex> lock(ctx->plan_mutex);
    for (const auto& sch : ctx->trsv_schedules)
      if (sch.analysis == analysis && sch.levels == info->levels) {
        lv.host_level_ptr = sch.level_ptr;
        break;
      }
  }
  const hipError_t e = asc_hip::trsv_solve_launch(
      rows, solve_lower, info->diagonal == ASC_HIP_UNIT, alpha.real,
      reinterpret_cast<const int64_t*>(w + L.level_ptr),
      reinterpret_cast<const uint32_t*>(w + L.row_order), lv, x, rstride, cstride, rhs->cols,
      ctx->stream);
  return finish_launch(ctx, e, "triangular solve");
}

}  // extern "C"
