// C ABI: sparse Level-1 over canonical indexed vectors (blas.h:1330-1582),
// replacing CloneIndexedErased / StandardLevel1Erased
// (include/asc/sparse/providers/cuda.h:216-218,234-238).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using namespace asc_hip_abi;

namespace {

bool host_space(uint8_t s) { return s == ASC_HIP_HOST || s == ASC_HIP_PINNED_HOST; }

asc_hip_status check_indexed_meta(const asc_hip_indexed_desc* d) {
  if (d == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "indexed vector descriptor is null");
  if (d->element_kind == ASC_HIP_C64 || d->element_kind == ASC_HIP_C128)
    return fail(ASC_HIP_UNSUPPORTED, "sparse Level-1 supports real FP32/FP64 elements");
  if (d->element_kind != ASC_HIP_F32 && d->element_kind != ASC_HIP_F64)
    return fail(ASC_HIP_INVALID_ARGUMENT, "unknown element kind");
  // blas.h:575-579
  if (d->nnz < 0 || d->dense_extent < 0 || d->nnz > d->dense_extent)
    return fail(ASC_HIP_INVALID_ARGUMENT, "A sparse indexed-vector descriptor is invalid");
  if (d->nnz > 0 && (d->indices == nullptr || d->values == nullptr))
    return fail(ASC_HIP_INVALID_ARGUMENT, "indexed vector arrays are null");
  return ok();
}

SparseSpans indexed_spans(const asc_hip_indexed_desc& d) {
  SparseSpans s;
  const size_t n = static_cast<size_t>(d.nnz);
  s.first = span_of(d.indices, 8 * n);
  s.values = span_of(d.values, elem_size(d.element_kind) * n);
  return s;
}

}  // namespace

extern "C" {

asc_hip_status asc_sparse_hip_clone_indexed(asc_hip_context* ctx, const asc_hip_indexed_desc* src,
                                            void* d_idx, void* d_val, asc_hip_indexed_desc* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  if ((st = check_indexed_meta(src)).code) return st;
  if (!host_space(src->memory_space)) return fail(ASC_HIP_MEMORY_ACCESS, "clone source must be host memory");
  // blas.h:606-620: sorted, unique, zero-based and in range.
  int64_t previous = -1;
  for (int64_t p = 0; p < src->nnz; ++p) {
    const int64_t i = src->indices[p];
    if (i <= previous || i < 0 || i >= src->dense_extent)
      return fail(ASC_HIP_INDEX,
                  "Sparse indexed-vector indices must be sorted, unique, zero-based, and in range");
    previous = i;
  }
  const size_t ib = 8 * static_cast<size_t>(src->nnz);
  const size_t vb = elem_size(src->element_kind) * static_cast<size_t>(src->nnz);
  if (src->nnz > 0) {
    if ((st = check_device_pointer(ctx, d_idx, "indices destination", ib)).code) return st;
    if ((st = check_device_pointer(ctx, d_val, "values destination", vb)).code) return st;
    if (overlaps(span_of(d_idx, ib), span_of(d_val, vb)))
      return fail(ASC_HIP_INVALID_ARGUMENT, "clone destinations overlap");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e = hipSuccess;
  if (src->nnz > 0) {
    e = hipMemcpyAsync(d_idx, src->indices, ib, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_val, src->values, vb, hipMemcpyHostToDevice, ctx->stream);
  }
  const hipError_t sync = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = sync;
  if (e != hipSuccess) return finish_launch(ctx, e, "indexed clone copy");
  *out = *src;
  out->indices = static_cast<const int64_t*>(d_idx);
  out->values = d_val;
  out->memory_space = ASC_HIP_DEVICE;
  out->writable = 1;
  out->trusted = 1;
  return ok();
}

asc_hip_status asc_sparse_hip_standard_level1(asc_hip_context* ctx, uint8_t operation,
                                              uint8_t conjugation, asc_hip_scalar alpha,
                                              const asc_hip_indexed_desc* sparse,
                                              const asc_hip_vector_desc* dense,
                                              const asc_hip_vector_desc* result,
                                              const asc_hip_spmv_options* options) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (operation > ASC_HIP_LEVEL1_SCATTER) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown Level-1 operation");
  if (conjugation > ASC_HIP_CONJUGATED) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown conjugation");
  if ((st = check_indexed_meta(sparse)).code) return st;
  if (sparse->memory_space != ASC_HIP_DEVICE && sparse->memory_space != ASC_HIP_MANAGED)
    return fail(ASC_HIP_MEMORY_ACCESS, "indexed vector is not device memory");
  if (!sparse->trusted)
    return fail(ASC_HIP_INVALID_ARGUMENT, "operations require a trusted canonical indexed vector");
  if (sparse->nnz > 0) {
    const size_t n = static_cast<size_t>(sparse->nnz);
    if ((st = check_device_pointer(ctx, sparse->indices, "indexed vector", 8 * n)).code) return st;
    if ((st = check_device_pointer(ctx, sparse->values, "indexed vector",
                                   elem_size(sparse->element_kind) * n)).code)
      return st;
  }
  const bool dot = operation == ASC_HIP_LEVEL1_DOT;
  const bool writes_dense = operation == ASC_HIP_LEVEL1_AXPY ||
                            operation == ASC_HIP_LEVEL1_GATHER_ZERO ||
                            operation == ASC_HIP_LEVEL1_SCATTER;
  const bool writes_sparse = operation == ASC_HIP_LEVEL1_GATHER ||
                             operation == ASC_HIP_LEVEL1_GATHER_ZERO;
  if (dense == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "dense descriptor is null");
  if (dense->element_kind != sparse->element_kind)
    return fail(ASC_HIP_INVALID_ARGUMENT, "dense element kind does not match the indexed vector");
  if (dense->extent != sparse->dense_extent)
    return fail(ASC_HIP_SHAPE, "Sparse Level-1 vector extents do not match");
  if ((st = check_vector(ctx, dense, sparse->element_kind, sparse->dense_extent, writes_dense,
                         "dense")).code)
    return st;
  if (writes_sparse && !sparse->writable)
    return fail(ASC_HIP_INVALID_ARGUMENT, "gathered values are not writable");
  const SparseSpans sp = indexed_spans(*sparse);
  const Span ds = vector_span(*dense);
  if ((writes_dense || writes_sparse) && (overlaps(ds, sp.first) || overlaps(ds, sp.values)))
    return fail(ASC_HIP_INVALID_ARGUMENT, "Sparse Level-1 writable operands overlap");
  if (dot) {
    if ((st = check_vector(ctx, result, sparse->element_kind, 1, true, "result")).code) return st;
    const Span rs = vector_span(*result);
    if (overlaps(rs, ds) || overlaps(rs, sp.first) || overlaps(rs, sp.values))
      return fail(ASC_HIP_INVALID_ARGUMENT, "dot result overlaps an operand");
  }
  const uint8_t alg = options ? options->algorithm : ASC_HIP_ALG_AUTO;
  if (alg != ASC_HIP_ALG_AUTO && alg != ASC_HIP_ALG_ORDERED)
    return fail(ASC_HIP_INVALID_ARGUMENT, "Level-1 algorithms are AUTO (plain) or ORDERED");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (operation == ASC_HIP_LEVEL1_AXPY && alpha.real == 0.0) return ok();  // blas.h:1388-1390
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  const hipError_t e = asc_hip::level1_launch(
      operation, alg == ASC_HIP_ALG_ORDERED, static_cast<asc_hip::Elem>(sparse->element_kind),
      alpha.real, sparse->indices, const_cast<void*>(sparse->values),
      const_cast<void*>(dense->data), dense->stride, sparse->nnz,
      dot ? const_cast<void*>(result->data) : nullptr, ctx->stream);
  return finish_launch(ctx, e, "Level-1 launch");
}

}  // extern "C"
