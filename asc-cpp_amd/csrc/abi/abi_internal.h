// Shared internals of the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "asc_sparse_hip.h"
#include "../kernels/launch.h"

constexpr uint64_t kContextMagic = 0x4153435f48495031ull;  // "ASC_HIP1"

struct asc_hip_context {
  uint64_t magic = kContextMagic;
  int device = -1;  // -1: validation-only context
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  // A second stream (default priority, library-owned) and fork/join events:
  // work of one call that can run beside the call's main kernels goes there.
  asc_hip::SideStream side{};
  // Device word a kernel sets when it gives up a bounded wait
  // (asc_hip::kFaultStreamWait); read and cleared by
  // asc_sparse_hip_context_synchronize, which then fails with
  // ASC_HIP_INTERNAL.
  uint32_t* fault_word = nullptr;
  // This device's copy of the allocation-free entry point's scratch array
  // (found at context creation; null for a validation-only context).
  void* noalloc_scratch = nullptr;
  uint8_t determinism = ASC_HIP_DETERMINISTIC;
  std::mutex plan_mutex;  // guards the tuned-plan cache below
  struct Plan {
    const void* offsets;
    const void* indices;
    int64_t rows, cols, nnz;
    uint8_t element_kind;
    uint8_t algorithm;
    asc_hip::PanelPlan panel;  // ASC_HIP_ALG_PANEL only
  };
  std::vector<Plan> plans;
  // Level bounds of this context's triangular analyses, by analysis buffer
  // (they only shape the solve's launches; the device copy is the truth).
  struct TrsvSchedule {
    const void* analysis;
    int64_t levels;
    std::vector<int64_t> level_ptr;
  };
  std::vector<TrsvSchedule> trsv_schedules;  // guarded by plan_mutex
};

namespace asc_hip_abi {

extern thread_local std::string g_message;

asc_hip_status ok();
asc_hip_status fail(int32_t code, const std::string& message);
asc_hip_status fail_hip(hipError_t e, const std::string& what);

struct DeviceGuard {
  int previous = -1;
  hipError_t error = hipSuccess;
  explicit DeviceGuard(int device);
  ~DeviceGuard();
};

struct Span {
  uintptr_t begin = 0;
  uintptr_t end = 0;
};
// The operand spans of one call, held inline so that validation allocates
// nothing on the success path (no call names more than kMax).
struct SpanList {
  static constexpr size_t kMax = 12;
  Span items[kMax];
  size_t count = 0;
  SpanList() = default;
  SpanList(std::initializer_list<Span> spans) {
    for (const Span& s : spans) push(s);
  }
  void push(Span s) {
    if (count == kMax) std::abort();  // a call with more operands than kMax: a bug here
    items[count++] = s;
  }
  const Span* begin() const { return items; }
  const Span* end() const { return items + count; }
};
bool overlaps(Span a, Span b);
Span span_of(const void* p, size_t bytes);
Span vector_span(const asc_hip_vector_desc& v);
Span matrix_span(const asc_hip_matrix_desc& m);
struct SparseSpans {
  Span first, second, values;
};
SparseSpans sparse_spans(const asc_hip_sparse_desc& d);
bool sparse_overlaps(const asc_hip_sparse_desc& d, Span out);
bool checked_mul(int64_t a, int64_t b, int64_t* out);
size_t elem_size(uint8_t kind);

asc_hip_status check_context(const asc_hip_context* ctx);
// With bytes > 0, [p, p + bytes) must also lie inside one device allocation.
asc_hip_status check_device_pointer(const asc_hip_context* ctx, const void* p,
                                    const char* what, size_t bytes = 0);
asc_hip_status check_sparse_meta(const asc_hip_sparse_desc* d, const char* what);
asc_hip_status check_device_sparse(const asc_hip_context* ctx,
                                   const asc_hip_sparse_desc* d, const char* what);
asc_hip_status check_vector(const asc_hip_context* ctx, const asc_hip_vector_desc* v,
                            uint8_t kind, int64_t extent, bool writable,
                            const char* what);
asc_hip_status check_matrix(const asc_hip_context* ctx, const asc_hip_matrix_desc* m,
                            uint8_t kind, int64_t rows, int64_t cols, bool writable,
                            const char* what);
asc_hip_status check_workspace(const asc_hip_context* ctx, void* ws, size_t ws_bytes,
                               size_t needed, const SpanList& operands);
asc_hip_status unavailable_if_validation_only(const asc_hip_context* ctx);
asc_hip_status finish_launch(asc_hip_context* ctx, hipError_t e, const char* what);
asc_hip::Compressed as_compressed(const asc_hip_sparse_desc& d, bool row_view_of_csc);

// Host-forced fault injection (test hooks, see asc_sparse_hip_test_inject).
bool test_hook_consume_allocation_failure();
bool test_hook_consume_event_failure();
bool test_hook_consume_stream_stall();
// Address of the current device's copy of the allocation-free scratch array
// (resolved once per device; call with `device` current).
asc_hip_status noalloc_scratch(int device, void** out);

}  // namespace asc_hip_abi
