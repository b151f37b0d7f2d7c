// C ABI of ASC::sparse_hip: validation, error mapping and dispatch.
//
// Follows the reference provider's contract (src/sparse/cuda/operations.cc):
// every operand is validated before the first enqueue (:349-398), device
// matrices must be trusted (:163-166), outputs may not overlap inputs
// (:382-396), workspaces are caller-owned and size-checked (:952-1002), HIP
// errors map to asc::ErrorCode (src/sparse/cuda/context.cc:15-35), and a
// failure after enqueue drains the stream and keeps the first error
// (:1026-1046).
//
// A context created for device -1 is validation-only: it runs every check
// that needs no GPU and answers ASC_HIP_UNAVAILABLE where work would be
// enqueued, so the rejection matrix is testable on a machine without a GPU.
#include "asc_sparse_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../kernels/launch.h"
#include "abi_internal.h"

using asc_hip::Elem;

struct asc_hip_event {
  hipEvent_t event = nullptr;
  int device = 0;
};

namespace asc_hip_abi {

thread_local std::string g_message;

asc_hip_status ok() { return {ASC_HIP_OK, ASC_HIP_PROVIDER_NONE, 0}; }

asc_hip_status fail(int32_t code, const std::string& message) {
  g_message = message;
  return {code, ASC_HIP_PROVIDER_NONE, 0};
}

int32_t map_hip(hipError_t e) {
  switch (e) {
    case hipErrorOutOfMemory: return ASC_HIP_ALLOCATION;
    case hipErrorNoDevice:
    case hipErrorInvalidDevice: return ASC_HIP_UNAVAILABLE;
    default: return ASC_HIP_PROVIDER;
  }
}

asc_hip_status fail_hip(hipError_t e, const std::string& what) {
  g_message = what + ": " + hipGetErrorString(e);
  return {map_hip(e), ASC_HIP_PROVIDER_HIP, static_cast<int64_t>(e)};
}

// Sets the context's device for the scope and restores the previous one.
DeviceGuard::DeviceGuard(int device) {
  if (hipGetDevice(&previous) != hipSuccess) previous = -1;
  if (previous != device) error = hipSetDevice(device);
}
DeviceGuard::~DeviceGuard() {
  int now = -1;
  if (previous >= 0 && hipGetDevice(&now) == hipSuccess && now != previous)
    (void)hipSetDevice(previous);
}

bool checked_mul(int64_t a, int64_t b, int64_t* out) {
  if (a < 0 || b < 0) return false;
  if (a != 0 && b > std::numeric_limits<int64_t>::max() / a) return false;
  *out = a * b;
  return true;
}

size_t elem_size(uint8_t kind) {
  switch (kind) {
    case ASC_HIP_F32: return 4;
    case ASC_HIP_F64: return 8;
    case ASC_HIP_C64: return 8;
    case ASC_HIP_C128: return 16;
    case ASC_HIP_BF16:
    case ASC_HIP_F16: return 2;
    default: return 0;
  }
}

bool overlaps(Span a, Span b) {
  if (a.begin == a.end || b.begin == b.end) return false;
  return a.begin < b.end && b.begin < a.end;
}

Span span_of(const void* p, size_t bytes) {
  Span s;
  s.begin = reinterpret_cast<uintptr_t>(p);
  s.end = s.begin + bytes;
  return s;
}

Span vector_span(const asc_hip_vector_desc& v) {
  Span s;
  if (v.extent == 0 || v.data == nullptr) return s;
  const int64_t esz = static_cast<int64_t>(elem_size(v.element_kind));
  const int64_t last = (v.extent - 1) * v.stride;
  const int64_t lo = std::min<int64_t>(0, last), hi = std::max<int64_t>(0, last);
  s.begin = reinterpret_cast<uintptr_t>(v.data) + lo * esz;
  s.end = reinterpret_cast<uintptr_t>(v.data) + hi * esz + esz;
  return s;
}

Span matrix_span(const asc_hip_matrix_desc& m) {
  Span s;
  if (m.rows == 0 || m.cols == 0 || m.data == nullptr) return s;
  const int64_t esz = static_cast<int64_t>(elem_size(m.element_kind));
  const int64_t major = m.layout == ASC_HIP_ROW_MAJOR ? m.rows : m.cols;
  const int64_t minor = m.layout == ASC_HIP_ROW_MAJOR ? m.cols : m.rows;
  s.begin = reinterpret_cast<uintptr_t>(m.data);
  s.end = s.begin + static_cast<uintptr_t>(((major - 1) * m.leading_dimension + minor) * esz);
  return s;
}

// Byte spans of a compressed/COO operand's three arrays.
SparseSpans sparse_spans(const asc_hip_sparse_desc& d) {
  SparseSpans s;
  const int64_t outer = d.format == ASC_HIP_CSC ? d.cols : d.rows;
  if (d.format == ASC_HIP_COO) {
    s.first = span_of(d.outer_offsets, static_cast<size_t>(d.nnz) * 16);
  } else {
    s.first = span_of(d.outer_offsets, static_cast<size_t>(outer + 1) * d.offset_bytes);
    s.second = span_of(d.inner_indices, static_cast<size_t>(d.nnz) * d.index_bytes);
  }
  s.values = span_of(d.values, static_cast<size_t>(d.nnz) * elem_size(d.element_kind));
  return s;
}

bool sparse_overlaps(const asc_hip_sparse_desc& d, Span out) {
  const SparseSpans s = sparse_spans(d);
  return overlaps(s.first, out) || overlaps(s.second, out) || overlaps(s.values, out);
}

// The pointer must be device (or managed) memory of the context's device;
// for device memory, bytes > 0 must also fit the allocation holding p
// (placement checked before enqueue: sparse.md:119-125, 455-466).
asc_hip_status check_device_pointer(const asc_hip_context* ctx, const void* p,
                                    const char* what, size_t bytes) {
  if (ctx->device < 0) return ok();  // validation-only context
  if (p == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + " is null");
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ASC_HIP_MEMORY_ACCESS,
                std::string(what) + " is not HIP device memory of this context");
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) {
    return fail(ASC_HIP_MEMORY_ACCESS,
                std::string(what) + " is not device-accessible memory");
  }
  if (attr.type == hipMemoryTypeDevice && attr.device != ctx->device) {
    return fail(ASC_HIP_MEMORY_ACCESS,
                std::string(what) + " belongs to another device");
  }
  if (attr.type == hipMemoryTypeDevice && bytes > 0) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    const hipError_t r = hipMemGetAddressRange(&base, &size, const_cast<void*>(p));
    if (r != hipSuccess) {
      (void)hipGetLastError();
      return fail(ASC_HIP_MEMORY_ACCESS,
                  std::string(what) + ": its allocation cannot be determined");
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base);
    const uintptr_t q = reinterpret_cast<uintptr_t>(p);
    if (q < b || bytes > size - (q - b))
      return fail(ASC_HIP_MEMORY_ACCESS,
                  std::string(what) + " extends past the end of its allocation");
  }
  return ok();
}

asc_hip_status check_context(const asc_hip_context* ctx) {
  if (ctx == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "context is null");
  if (ctx->magic != kContextMagic)
    return fail(ASC_HIP_INVALID_STATE, "context is not a live asc_sparse_hip context");
  return ok();
}

// Shape/metadata checks shared by host and device sparse operands
// (operations.cc:134-220).
asc_hip_status check_sparse_meta(const asc_hip_sparse_desc* d, const char* what) {
  if (d == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + " descriptor is null");
  if (d->format > ASC_HIP_CSC) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": unknown format");
  if (d->rows < 0 || d->cols < 0) return fail(ASC_HIP_SHAPE, std::string(what) + ": negative extent");
  int64_t logical = 0;
  if (!checked_mul(d->rows, d->cols, &logical)) return fail(ASC_HIP_OVERFLOW, std::string(what) + ": extent product overflows");
  if (d->nnz < 0 || d->nnz > logical) return fail(ASC_HIP_SHAPE, std::string(what) + ": nnz is incompatible with its shape");
  if (elem_size(d->element_kind) == 0) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": unknown element kind");
  if (d->format != ASC_HIP_COO) {
    if (d->index_bytes != 4 && d->index_bytes != 8) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": index width must be 4 or 8 bytes");
    if (d->offset_bytes != 8) return fail(ASC_HIP_UNSUPPORTED, std::string(what) + ": offsets must be int64");
    const int64_t inner = d->format == ASC_HIP_CSR ? d->cols : d->rows;
    if (d->index_bytes == 4 && inner > std::numeric_limits<int32_t>::max())
      return fail(ASC_HIP_OVERFLOW, std::string(what) + ": int32 indices cannot address the inner extent");
    if (d->outer_offsets == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": offsets are null");
  } else if (d->nnz > 0 && d->outer_offsets == nullptr) {
    return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": coordinates are null");
  }
  if (d->nnz > 0 && (d->values == nullptr || (d->format != ASC_HIP_COO && d->inner_indices == nullptr)))
    return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": arrays are null");
  // Every array's byte count must be representable (CheckedByteCount,
  // compressed.h:477-497): the spans checked against allocations use them.
  const int64_t outer = d->format == ASC_HIP_CSC ? d->cols : d->rows;
  int64_t bytes = 0;
  const bool sized =
      d->format == ASC_HIP_COO
          ? checked_mul(d->nnz, 16, &bytes)
          : outer < std::numeric_limits<int64_t>::max() &&
                checked_mul(outer + 1, d->offset_bytes, &bytes) &&
                checked_mul(d->nnz, d->index_bytes, &bytes);
  if (!sized || !checked_mul(d->nnz, static_cast<int64_t>(elem_size(d->element_kind)), &bytes))
    return fail(ASC_HIP_OVERFLOW, std::string(what) + ": array byte count overflows");
  return ok();
}

asc_hip_status check_device_sparse(const asc_hip_context* ctx, const asc_hip_sparse_desc* d,
                                   const char* what) {
  asc_hip_status st = check_sparse_meta(d, what);
  if (st.code) return st;
  if (d->memory_space != ASC_HIP_DEVICE && d->memory_space != ASC_HIP_MANAGED)
    return fail(ASC_HIP_MEMORY_ACCESS, std::string(what) + ": operand is not device memory");
  if (!d->trusted)
    return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": operations require trusted canonical structure");
  const SparseSpans s = sparse_spans(*d);
  if ((st = check_device_pointer(ctx, d->outer_offsets, what, s.first.end - s.first.begin)).code)
    return st;
  if (d->nnz > 0) {
    if (d->format != ASC_HIP_COO &&
        (st = check_device_pointer(ctx, d->inner_indices, what, s.second.end - s.second.begin)).code)
      return st;
    if ((st = check_device_pointer(ctx, d->values, what, s.values.end - s.values.begin)).code)
      return st;
  }
  return ok();
}

asc_hip_status check_vector(const asc_hip_context* ctx, const asc_hip_vector_desc* v,
                            uint8_t kind, int64_t extent, bool writable, const char* what) {
  if (v == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + " descriptor is null");
  if (v->element_kind != kind) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": element kind does not match the operation");
  if (v->extent != extent) return fail(ASC_HIP_SHAPE, std::string(what) + ": length does not match the matrix");
  if (v->stride == 0 && extent > 1) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": stride must be nonzero");
  {
    // The span (extent - 1) * |stride| + 1 elements must have a byte count.
    const int64_t mag = v->stride == std::numeric_limits<int64_t>::min()
                            ? -1
                            : (v->stride < 0 ? -v->stride : v->stride);
    int64_t reach = 0, bytes = 0;
    if (extent > 1 && (mag < 0 || !checked_mul(extent - 1, mag, &reach) ||
                       reach == std::numeric_limits<int64_t>::max() ||
                       !checked_mul(reach + 1, static_cast<int64_t>(elem_size(kind)), &bytes)))
      return fail(ASC_HIP_OVERFLOW, std::string(what) + ": strided span overflows");
  }
  if (writable && !v->writable) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": output is not writable");
  if (v->memory_space != ASC_HIP_DEVICE && v->memory_space != ASC_HIP_MANAGED)
    return fail(ASC_HIP_MEMORY_ACCESS, std::string(what) + ": vector is not device memory");
  if (extent > 0) {
    asc_hip_status st = check_device_pointer(ctx, v->data, what);
    if (st.code) return st;
    const Span s = vector_span(*v);
    st = check_device_pointer(ctx, reinterpret_cast<const void*>(s.begin), what, s.end - s.begin);
    if (st.code) return st;
  }
  return ok();
}

asc_hip_status check_workspace(const asc_hip_context* ctx, void* ws, size_t ws_bytes,
                               size_t needed, const SpanList& operands) {
  if (needed == 0) return ok();
  if (ws == nullptr || ws_bytes < needed)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "workspace is smaller than the queried size (" + std::to_string(needed) + " bytes)");
  if (reinterpret_cast<uintptr_t>(ws) % 256 != 0)
    return fail(ASC_HIP_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
  const Span w = span_of(ws, needed);
  for (const Span& s : operands)
    if (overlaps(w, s)) return fail(ASC_HIP_INVALID_ARGUMENT, "workspace overlaps an operand");
  return check_device_pointer(ctx, ws, "workspace", needed);
}

asc_hip_status unavailable_if_validation_only(const asc_hip_context* ctx) {
  if (ctx->device < 0)
    return fail(ASC_HIP_UNAVAILABLE, "validation-only context cannot enqueue work");
  return ok();
}

// Drains the stream after a failed enqueue and keeps the first error.
asc_hip_status finish_launch(asc_hip_context* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return ok();
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side.stream) (void)hipStreamSynchronize(ctx->side.stream);
  (void)hipGetLastError();
  return fail_hip(e, what);
}

asc_hip_status check_matrix(const asc_hip_context* ctx, const asc_hip_matrix_desc* m,
                            uint8_t kind, int64_t rows, int64_t cols, bool writable,
                            const char* what) {
  if (m == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + " descriptor is null");
  if (m->element_kind != kind) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": element kind does not match");
  if (m->layout > ASC_HIP_ROW_MAJOR) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": unknown layout");
  if (m->rows != rows || m->cols < 0 || (cols >= 0 && m->cols != cols))
    return fail(ASC_HIP_SHAPE, "Sparse matrix-matrix operand extents do not match");
  const int64_t minor = m->layout == ASC_HIP_ROW_MAJOR ? m->cols : m->rows;
  if (m->leading_dimension < (minor > 1 ? minor : 1))
    return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": leading dimension is too small");
  if (m->rows > 0 && m->cols > 0) {
    // ((major - 1) * ld + minor) elements must have a byte count.
    const int64_t major = m->layout == ASC_HIP_ROW_MAJOR ? m->rows : m->cols;
    int64_t reach = 0, bytes = 0;
    if (!checked_mul(major - 1, m->leading_dimension, &reach) ||
        reach > std::numeric_limits<int64_t>::max() - minor ||
        !checked_mul(reach + minor, static_cast<int64_t>(elem_size(kind)), &bytes))
      return fail(ASC_HIP_OVERFLOW, std::string(what) + ": block span overflows");
  }
  if (writable && !m->writable) return fail(ASC_HIP_INVALID_ARGUMENT, std::string(what) + ": output is not writable");
  if (m->memory_space != ASC_HIP_DEVICE && m->memory_space != ASC_HIP_MANAGED)
    return fail(ASC_HIP_MEMORY_ACCESS, std::string(what) + ": block is not device memory");
  if (m->rows > 0 && m->cols > 0) {
    const Span s = matrix_span(*m);
    return check_device_pointer(ctx, m->data, what, s.end - s.begin);
  }
  return ok();
}

asc_hip::Compressed as_compressed(const asc_hip_sparse_desc& d, bool row_view_of_csc) {
  asc_hip::Compressed c;
  c.offsets = static_cast<const int64_t*>(d.outer_offsets);
  c.indices = d.inner_indices;
  c.values = d.values;
  const bool csr = d.format == ASC_HIP_CSR;
  (void)row_view_of_csc;
  c.outer = csr ? d.rows : d.cols;
  c.inner = csr ? d.cols : d.rows;
  c.nnz = d.nnz;
  c.index_bytes = d.index_bytes;
  c.elem = static_cast<Elem>(d.element_kind);
  return c;
}

std::atomic<int32_t> g_fail_allocation{-1};
std::atomic<int32_t> g_fail_event{-1};
std::atomic<int32_t> g_stream_stall{-1};

bool consume(std::atomic<int32_t>& countdown) {
  int32_t v = countdown.load();
  while (v >= 0) {
    if (countdown.compare_exchange_weak(v, v - 1)) return v == 0;
  }
  return false;
}

bool test_hook_consume_allocation_failure() { return consume(g_fail_allocation); }
bool test_hook_consume_event_failure() { return consume(g_fail_event); }
bool test_hook_consume_stream_stall() { return consume(g_stream_stall); }

}  // namespace asc_hip_abi

using namespace asc_hip_abi;

extern "C" {

const char* asc_sparse_hip_last_message(void) { return g_message.c_str(); }

asc_hip_status asc_sparse_hip_test_inject(uint32_t what, int32_t countdown) {
  if (what == ASC_HIP_INJECT_ALLOCATION) g_fail_allocation.store(countdown);
  else if (what == ASC_HIP_INJECT_EVENT) g_fail_event.store(countdown);
  else if (what == ASC_HIP_INJECT_STREAM_STALL) g_stream_stall.store(countdown);
  else return fail(ASC_HIP_INVALID_ARGUMENT, "unknown injection point");
  return ok();
}

int32_t asc_sparse_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

asc_hip_status asc_sparse_hip_context_create(int32_t device, void* stream,
                                             uint8_t determinism,
                                             asc_hip_context** out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = nullptr;
  if (determinism > ASC_HIP_BACKEND_DEFAULT) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown determinism");
  auto* ctx = new (std::nothrow) asc_hip_context();
  if (ctx == nullptr) return fail(ASC_HIP_ALLOCATION, "context allocation failed");
  ctx->determinism = determinism;
  if (device == -1) {  // validation-only
    ctx->device = -1;
    *out = ctx;
    return ok();
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    delete ctx;
    return fail(ASC_HIP_UNAVAILABLE, "no HIP device is available");
  }
  if (device < 0 || device >= count) {
    delete ctx;
    return fail(ASC_HIP_UNAVAILABLE, "device ordinal out of range");
  }
  ctx->device = device;
  DeviceGuard guard(device);
  if (guard.error != hipSuccess) {
    delete ctx;
    return fail_hip(guard.error, "hipSetDevice");
  }
  if (stream != nullptr) {
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->owns_stream = false;
  } else {
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      return fail_hip(e, "hipStreamCreateWithFlags");
    }
    ctx->owns_stream = true;
  }
  e = hipStreamCreateWithFlags(&ctx->side.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->side.fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->side.join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&ctx->fault_word), sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(ctx->fault_word, 0, sizeof(uint32_t));
  if (e == hipSuccess) {
    const asc_hip_status sst = noalloc_scratch(device, &ctx->noalloc_scratch);
    if (sst.code != ASC_HIP_OK) {
      (void)hipFree(ctx->fault_word);
      (void)hipEventDestroy(ctx->side.join);
      (void)hipEventDestroy(ctx->side.fork);
      (void)hipStreamDestroy(ctx->side.stream);
      if (ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
      delete ctx;
      return sst;
    }
  }
  if (e != hipSuccess) {
    if (ctx->fault_word) (void)hipFree(ctx->fault_word);
    if (ctx->side.join) (void)hipEventDestroy(ctx->side.join);
    if (ctx->side.fork) (void)hipEventDestroy(ctx->side.fork);
    if (ctx->side.stream) (void)hipStreamDestroy(ctx->side.stream);
    if (ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return fail_hip(e, "creating the side stream and fault word");
  }
  *out = ctx;
  return ok();
}

asc_hip_status asc_sparse_hip_context_destroy(asc_hip_context* ctx) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (ctx->device >= 0) {
    DeviceGuard guard(ctx->device);
    if (ctx->owns_stream) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->side.stream) {
      (void)hipStreamSynchronize(ctx->side.stream);
      (void)hipStreamDestroy(ctx->side.stream);
    }
    if (ctx->side.fork) (void)hipEventDestroy(ctx->side.fork);
    if (ctx->side.join) (void)hipEventDestroy(ctx->side.join);
    if (ctx->fault_word) (void)hipFree(ctx->fault_word);
  }
  ctx->magic = 0;
  delete ctx;
  return ok();
}

asc_hip_status asc_sparse_hip_context_stream(asc_hip_context* ctx, void** out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = ctx->stream;
  return ok();
}

asc_hip_status asc_sparse_hip_context_synchronize(asc_hip_context* ctx) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
  // A kernel that gave up a bounded wait left its mark: report it once.
  uint32_t fault = 0;
  e = hipMemcpy(&fault, ctx->fault_word, sizeof(fault), hipMemcpyDeviceToHost);
  if (e == hipSuccess && fault != 0) e = hipMemset(ctx->fault_word, 0, sizeof(uint32_t));
  if (e != hipSuccess) return fail_hip(e, "reading the context's fault word");
  if (fault & asc_hip::kFaultStreamWait)
    return fail(ASC_HIP_INTERNAL,
                "a stream-staged panel SpMV gave up waiting between its waves; the outputs of "
                "the SpMVs since the last synchronization are not valid");
  return ok();
}

asc_hip_status asc_sparse_hip_allocate(asc_hip_context* ctx, size_t bytes,
                                       uint8_t space, void** out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = nullptr;
  if (space > ASC_HIP_MANAGED) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown memory space");
  if (bytes == 0) return ok();
  if (test_hook_consume_allocation_failure())
    return fail(ASC_HIP_ALLOCATION, "injected allocation failure");
  if (space == ASC_HIP_HOST) {
    void* p = std::aligned_alloc(256, (bytes + 255) & ~size_t(255));
    if (p == nullptr) return fail(ASC_HIP_ALLOCATION, "host allocation failed");
    *out = p;
    return ok();
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e;
  if (space == ASC_HIP_DEVICE) e = hipMalloc(out, bytes);
  else if (space == ASC_HIP_PINNED_HOST) e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  else e = hipMallocManaged(out, bytes, hipMemAttachGlobal);
  if (e != hipSuccess) {
    *out = nullptr;
    (void)hipGetLastError();
    return fail_hip(e, "device allocation");
  }
  return ok();
}

asc_hip_status asc_sparse_hip_deallocate(asc_hip_context* ctx, void* ptr, uint8_t space) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (ptr == nullptr) return ok();
  if (space == ASC_HIP_HOST) {
    std::free(ptr);
    return ok();
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e;
  if (space == ASC_HIP_PINNED_HOST) e = hipHostFree(ptr);
  else e = hipFree(ptr);
  return e == hipSuccess ? ok() : fail_hip(e, "device free");
}

asc_hip_status asc_sparse_hip_copy(asc_hip_context* ctx, void* dst, const void* src,
                                   size_t bytes) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (bytes == 0) return ok();
  if (dst == nullptr || src == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "copy pointer is null");
  if (overlaps(span_of(dst, bytes), span_of(src, bytes)))
    return fail(ASC_HIP_INVALID_ARGUMENT, "copy source and destination overlap");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream);
  if (e != hipSuccess) return finish_launch(ctx, e, "hipMemcpyAsync");
  return ok();
}

asc_hip_status asc_sparse_hip_fill(asc_hip_context* ctx, void* dst, uint8_t byte, size_t bytes) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (bytes == 0) return ok();
  if ((st = check_device_pointer(ctx, dst, "fill destination")).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::fill_launch(dst, byte, bytes, ctx->stream), "fill");
}

asc_hip_status asc_sparse_hip_event_record(asc_hip_context* ctx, asc_hip_event** out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = nullptr;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (test_hook_consume_event_failure())
    return fail(ASC_HIP_PROVIDER, "injected event failure");
  DeviceGuard guard(ctx->device);
  auto* ev = new (std::nothrow) asc_hip_event();
  if (ev == nullptr) return fail(ASC_HIP_ALLOCATION, "event allocation failed");
  ev->device = ctx->device;
  hipError_t e = hipEventCreate(&ev->event);
  if (e != hipSuccess) {
    delete ev;
    return fail_hip(e, "hipEventCreate");
  }
  e = hipEventRecord(ev->event, ctx->stream);
  if (e != hipSuccess) {
    (void)hipEventDestroy(ev->event);
    delete ev;
    return fail_hip(e, "hipEventRecord");
  }
  *out = ev;
  return ok();
}

asc_hip_status asc_sparse_hip_event_wait(asc_hip_event* ev) {
  if (ev == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "event is null");
  DeviceGuard guard(ev->device);
  const hipError_t e = hipEventSynchronize(ev->event);
  return e == hipSuccess ? ok() : fail_hip(e, "hipEventSynchronize");
}

asc_hip_status asc_sparse_hip_event_elapsed_ms(asc_hip_event* a, asc_hip_event* b, float* out) {
  if (a == nullptr || b == nullptr || out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "null argument");
  DeviceGuard guard(a->device);
  const hipError_t e = hipEventElapsedTime(out, a->event, b->event);
  return e == hipSuccess ? ok() : fail_hip(e, "hipEventElapsedTime");
}

asc_hip_status asc_sparse_hip_event_destroy(asc_hip_event* ev) {
  if (ev == nullptr) return ok();
  DeviceGuard guard(ev->device);
  const hipError_t e = hipEventDestroy(ev->event);
  delete ev;
  return e == hipSuccess ? ok() : fail_hip(e, "hipEventDestroy");
}

}  // extern "C"
