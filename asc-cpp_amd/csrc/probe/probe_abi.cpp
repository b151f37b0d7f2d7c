// C ABI of the measurement probes (include/asc_sparse_hip_probe.h; no
// reference counterpart).  Validated like the library's operations: device
// pointers, spans inside their allocations, no overlap of an output with the
// inputs.  Links against libasc_sparse_hip.so for the context and checks.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../abi/abi_internal.h"
#include "asc_sparse_hip_probe.h"
#include "probe_launch.h"

using namespace asc_hip_abi;

static_assert(ASC_HIP_GATHER_PROBE_BLOCKS == asc_hip::kGatherProbeBlocks, "probe grid");
static_assert(ASC_HIP_ROW_GATHER_PROBE_BLOCKS == asc_hip::kRowGatherProbeBlocks, "probe grid");
static_assert(ASC_HIP_STREAM_PROBE_BLOCKS == asc_hip::kStreamProbeBlocks, "probe grid");

extern "C" {

asc_hip_status asc_sparse_hip_gather_probe(asc_hip_context* ctx, const int32_t* indices,
                                           const double* values, int64_t count, const double* x,
                                           int64_t x_extent, double* partials,
                                           size_t partials_count) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (count < 0 || x_extent < 0) return fail(ASC_HIP_SHAPE, "negative probe extent");
  if (x_extent > static_cast<int64_t>(UINT32_MAX) + 1)
    return fail(ASC_HIP_UNSUPPORTED, "the probe gathers from at most 2^32 elements");
  if (count > 0 && x_extent == 0) return fail(ASC_HIP_SHAPE, "gathers need a nonempty x");
  if (partials == nullptr || partials_count < static_cast<size_t>(asc_hip::kGatherProbeBlocks))
    return fail(ASC_HIP_INVALID_ARGUMENT, "partials must hold ASC_HIP_GATHER_PROBE_BLOCKS values");
  const size_t pb = sizeof(double) * static_cast<size_t>(asc_hip::kGatherProbeBlocks);
  if ((st = check_device_pointer(ctx, partials, "partials", pb)).code) return st;
  const size_t n = static_cast<size_t>(count);
  if (count > 0) {
    if ((st = check_device_pointer(ctx, indices, "indices", 4 * n)).code) return st;
    if ((st = check_device_pointer(ctx, values, "values", 8 * n)).code) return st;
    if ((st = check_device_pointer(ctx, x, "x", 8 * static_cast<size_t>(x_extent))).code)
      return st;
    const Span out = span_of(partials, pb);
    if (overlaps(out, span_of(indices, 4 * n)) || overlaps(out, span_of(values, 8 * n)) ||
        overlaps(out, span_of(x, 8 * static_cast<size_t>(x_extent))))
      return fail(ASC_HIP_INVALID_ARGUMENT, "partials overlap an input");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx,
                       asc_hip::gather_probe_launch(indices, values, count, x, x_extent, partials,
                                                    ctx->stream),
                       "gather probe");
}

asc_hip_status asc_sparse_hip_row_gather_probe(asc_hip_context* ctx, const int32_t* indices,
                                               int64_t count, const double* table,
                                               int64_t table_rows, double* partials,
                                               size_t partials_count) {
  constexpr size_t kRowBytes = 256;
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (count < 0 || table_rows < 0) return fail(ASC_HIP_SHAPE, "negative probe extent");
  if (table_rows > static_cast<int64_t>(UINT32_MAX) + 1)
    return fail(ASC_HIP_UNSUPPORTED, "the probe gathers from at most 2^32 rows");
  if (count > 0 && table_rows == 0) return fail(ASC_HIP_SHAPE, "gathers need a nonempty table");
  if (partials == nullptr ||
      partials_count < static_cast<size_t>(asc_hip::kRowGatherProbeBlocks))
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "partials must hold ASC_HIP_ROW_GATHER_PROBE_BLOCKS values");
  const size_t pb = sizeof(double) * static_cast<size_t>(asc_hip::kRowGatherProbeBlocks);
  if ((st = check_device_pointer(ctx, partials, "partials", pb)).code) return st;
  const size_t n = static_cast<size_t>(count);
  if (count > 0) {
    if (reinterpret_cast<uintptr_t>(table) % 32 != 0)
      return fail(ASC_HIP_INVALID_ARGUMENT, "the table must be 32-byte aligned");
    const size_t tb = kRowBytes * static_cast<size_t>(table_rows);
    if ((st = check_device_pointer(ctx, indices, "indices", 4 * n)).code) return st;
    if ((st = check_device_pointer(ctx, table, "table", tb)).code) return st;
    const Span out = span_of(partials, pb);
    if (overlaps(out, span_of(indices, 4 * n)) || overlaps(out, span_of(table, tb)))
      return fail(ASC_HIP_INVALID_ARGUMENT, "partials overlap an input");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx,
                       asc_hip::row_gather_probe_launch(indices, count, table, table_rows,
                                                        partials, ctx->stream),
                       "row gather probe");
}

asc_hip_status asc_sparse_hip_stream_probe(asc_hip_context* ctx, const void* src, void* dst,
                                           int64_t bytes, uint32_t* partials,
                                           size_t partials_count) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (bytes < 0 || bytes % 16 != 0)
    return fail(ASC_HIP_SHAPE, "the stream probe moves a nonnegative multiple of 16 bytes");
  const size_t n = static_cast<size_t>(bytes);
  if (bytes > 0) {
    if (reinterpret_cast<uintptr_t>(src) % 16 != 0 || reinterpret_cast<uintptr_t>(dst) % 16 != 0)
      return fail(ASC_HIP_INVALID_ARGUMENT, "stream probe buffers must be 16-byte aligned");
    if ((st = check_device_pointer(ctx, src, "source", n)).code) return st;
    if (dst != nullptr) {
      if ((st = check_device_pointer(ctx, dst, "destination", n)).code) return st;
      if (overlaps(span_of(dst, n), span_of(src, n)))
        return fail(ASC_HIP_INVALID_ARGUMENT, "destination overlaps the source");
    } else {
      if (partials == nullptr || partials_count < static_cast<size_t>(asc_hip::kStreamProbeBlocks))
        return fail(ASC_HIP_INVALID_ARGUMENT,
                    "partials must hold ASC_HIP_STREAM_PROBE_BLOCKS values");
      const size_t pb = sizeof(uint32_t) * static_cast<size_t>(asc_hip::kStreamProbeBlocks);
      if ((st = check_device_pointer(ctx, partials, "partials", pb)).code) return st;
      if (overlaps(span_of(partials, pb), span_of(src, n)))
        return fail(ASC_HIP_INVALID_ARGUMENT, "partials overlap the source");
    }
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::stream_probe_launch(src, dst, bytes / 16, partials, ctx->stream),
                       "stream probe");
}

}  // extern "C"
