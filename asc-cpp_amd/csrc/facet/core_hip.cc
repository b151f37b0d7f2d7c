// ASC::core_hip (include/asc_sparse_hip/core_hip.hpp): HIP device memory,
// execution contexts and completion events on the reference's core seam.
//
// Follows the contract of the reference's CUDA core provider
// (src/core/cuda/runtime.cc): a context owns one non-blocking stream; copies
// validate both pointers' placement against the declared memory spaces, reject
// partial overlap and record a completion event; a failure after enqueue
// drains the stream before it is reported; HIP errors map to ErrorCode with
// provider "hip" and the hipError_t as native code.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <utility>

#include "asc_sparse_hip/core_hip.hpp"
#if defined(ASC_SPARSE_HIP_IN_ASC_TREE)
#include "execution_internal.h"  // src/core/execution_internal.h of the asc-cpp tree
#endif

namespace asc {
namespace internal_core_hip {
namespace {

Status HipStatus(hipError_t error, ErrorCode fallback, const char* message) {
  ErrorCode code = fallback;
  if (error == hipErrorOutOfMemory) code = ErrorCode::kAllocation;
  if (error == hipErrorNoDevice || error == hipErrorInvalidDevice) code = ErrorCode::kUnavailable;
  (void)hipGetLastError();  // clear the sticky non-fatal error
  return Status(code, message, "hip", static_cast<std::int64_t>(error));
}

// Makes `device` current for a scope and restores the previous device.
class DeviceGuard {
 public:
  explicit DeviceGuard(std::int32_t device) {
    error_ = hipGetDevice(&previous_);
    if (error_ == hipSuccess && previous_ != device) {
      error_ = hipSetDevice(device);
      restore_ = error_ == hipSuccess;
    }
  }
  ~DeviceGuard() {
    if (restore_) (void)hipSetDevice(previous_);
  }
  DeviceGuard(const DeviceGuard&) = delete;
  DeviceGuard& operator=(const DeviceGuard&) = delete;
  [[nodiscard]] Status status() const {
    return error_ == hipSuccess ? Status()
                                : HipStatus(error_, ErrorCode::kUnavailable,
                                            "HIP could not select the context's device");
  }

 private:
  int previous_ = 0;
  bool restore_ = false;
  hipError_t error_ = hipSuccess;
};

Status ValidateDevice(std::int32_t device) {
  if (device < 0) return Status(ErrorCode::kInvalidArgument, "A HIP device ordinal cannot be negative");
  auto count = HipDeviceCount();
  if (!count.ok()) return count.status();
  if (device >= *count)
    return Status(ErrorCode::kUnavailable, "The requested HIP device is unavailable", "hip");
  return Status();
}

// The declared placement of a view against what HIP knows of the pointer.
Status ValidatePointer(const void* pointer, MemorySpace declared, std::int32_t device) {
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, pointer);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (declared == MemorySpace::kHost) return Status();  // plain host memory is unknown to HIP
    return Status(ErrorCode::kMemoryAccess, "HIP could not inspect a memory pointer", "hip",
                  static_cast<std::int64_t>(e));
  }
  switch (declared) {
    case MemorySpace::kHost:
      if (a.type != hipMemoryTypeUnregistered && a.type != hipMemoryTypeHost)
        return Status(ErrorCode::kMemoryAccess, "A pointer declared as host memory is HIP device memory");
      return Status();
    case MemorySpace::kPinnedHost:
      if (a.type != hipMemoryTypeHost)
        return Status(ErrorCode::kMemoryAccess, "A pinned-host view is not HIP-pinned host memory");
      return Status();
    case MemorySpace::kDevice:
      if (a.type != hipMemoryTypeDevice || a.device != device)
        return Status(ErrorCode::kMemoryAccess, "A device view is not on the execution-context device");
      return Status();
    case MemorySpace::kManaged:
      if (a.type != hipMemoryTypeManaged && a.type != hipMemoryTypeUnified)
        return Status(ErrorCode::kMemoryAccess, "A managed view is not HIP managed memory");
      return Status();
  }
  return Status(ErrorCode::kInvalidArgument, "A memory view has an unrecognized memory space");
}

bool RangesOverlap(const void* a, const void* b, std::size_t bytes) noexcept {
  const auto x = reinterpret_cast<std::uintptr_t>(a), y = reinterpret_cast<std::uintptr_t>(b);
  if (x > UINTPTR_MAX - bytes || y > UINTPTR_MAX - bytes) return true;
  return x < y + bytes && y < x + bytes;
}

class HipExecutionState;

class HipCompletionState final : public internal_core_execution::CompletionState {
 public:
  HipCompletionState(std::shared_ptr<const HipExecutionState> owner, hipEvent_t event) noexcept
      : owner_(std::move(owner)), event_(event) {}
  ~HipCompletionState() override;
  [[nodiscard]] Result<bool> Query() const override;
  [[nodiscard]] Status Wait() const override;

 private:
  std::shared_ptr<const HipExecutionState> owner_;  // keeps the stream's device alive
  hipEvent_t event_;
};

class HipExecutionState final : public internal_core_execution::ExecutionState,
                                public std::enable_shared_from_this<HipExecutionState> {
 public:
  HipExecutionState(std::int32_t device, hipStream_t stream) noexcept
      : device_(device), stream_(stream) {}
  ~HipExecutionState() override {
    DeviceGuard guard(device_);
    if (guard.status().ok()) (void)hipStreamDestroy(stream_);
  }
  [[nodiscard]] bool CanAccess(MemorySpace space) const noexcept override {
    switch (space) {
      case MemorySpace::kHost:
      case MemorySpace::kPinnedHost:
      case MemorySpace::kDevice:
      case MemorySpace::kManaged:
        return true;
    }
    return false;
  }
  [[nodiscard]] Result<std::unique_ptr<internal_core_execution::CompletionState>> CopyBytes(
      MutableMemoryView destination, ConstMemoryView source, std::size_t bytes) const override {
    DeviceGuard guard(device_);
    if (Status s = guard.status(); !s.ok()) return s;
    if (Status s = ValidatePointer(source.data(), source.space(), device_); !s.ok()) return s;
    if (Status s = ValidatePointer(destination.data(), destination.space(), device_); !s.ok())
      return s;
    if (source.data() == destination.data() && source.space() == destination.space())
      return RecordEvent();
    if (RangesOverlap(source.data(), destination.data(), bytes))
      return Status(ErrorCode::kInvalidArgument, "HIP CopyBytes rejects partially overlapping views");
    hipEvent_t event = nullptr;
    hipError_t e = hipEventCreateWithFlags(&event, hipEventDisableTiming);
    if (e != hipSuccess)
      return HipStatus(e, ErrorCode::kProvider, "HIP could not create a completion event");
    e = hipMemcpyAsync(destination.data(), source.data(), bytes, hipMemcpyDefault, stream_);
    if (e == hipSuccess) e = hipEventRecord(event, stream_);
    if (e != hipSuccess) {
      const Status failure =
          HipStatus(e, ErrorCode::kMemoryTransfer, "HIP could not enqueue an asynchronous copy");
      (void)hipStreamSynchronize(stream_);
      (void)hipEventDestroy(event);
      return failure;
    }
    return std::unique_ptr<internal_core_execution::CompletionState>(
        new HipCompletionState(shared_from_this(), event));
  }
  [[nodiscard]] Result<std::unique_ptr<internal_core_execution::CompletionState>> RecordEvent()
      const override {
    DeviceGuard guard(device_);
    if (Status s = guard.status(); !s.ok()) return s;
    hipEvent_t event = nullptr;
    hipError_t e = hipEventCreateWithFlags(&event, hipEventDisableTiming);
    if (e != hipSuccess) {
      const Status failure = HipStatus(e, ErrorCode::kProvider, "HIP could not create a completion event");
      (void)hipStreamSynchronize(stream_);
      return failure;
    }
    e = hipEventRecord(event, stream_);
    if (e != hipSuccess) {
      const Status failure = HipStatus(e, ErrorCode::kProvider, "HIP could not record a completion event");
      (void)hipStreamSynchronize(stream_);
      (void)hipEventDestroy(event);
      return failure;
    }
    return std::unique_ptr<internal_core_execution::CompletionState>(
        new HipCompletionState(shared_from_this(), event));
  }
  [[nodiscard]] void* NativeExecutionHandle(Backend backend) const noexcept override {
    return backend == Backend::kHip ? static_cast<void*>(stream_) : nullptr;
  }
  [[nodiscard]] std::int32_t device() const noexcept { return device_; }

 private:
  std::int32_t device_;
  hipStream_t stream_;
};

HipCompletionState::~HipCompletionState() {
  DeviceGuard guard(owner_->device());
  if (guard.status().ok()) (void)hipEventDestroy(event_);
}

Result<bool> HipCompletionState::Query() const {
  DeviceGuard guard(owner_->device());
  if (Status s = guard.status(); !s.ok()) return s;
  const hipError_t e = hipEventQuery(event_);
  if (e == hipSuccess) return true;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    return false;
  }
  return HipStatus(e, ErrorCode::kProvider, "HIP could not query a completion event");
}

Status HipCompletionState::Wait() const {
  DeviceGuard guard(owner_->device());
  if (Status s = guard.status(); !s.ok()) return s;
  const hipError_t e = hipEventSynchronize(event_);
  if (e != hipSuccess) return HipStatus(e, ErrorCode::kProvider, "HIP could not wait for a completion event");
  return Status::Ok();
}

}  // namespace
}  // namespace internal_core_hip

Result<std::int32_t> HipDeviceCount() {
  int count = 0;
  const hipError_t e = hipGetDeviceCount(&count);
  if (e == hipErrorNoDevice) {
    (void)hipGetLastError();
    return 0;
  }
  if (e != hipSuccess)
    return internal_core_hip::HipStatus(e, ErrorCode::kUnavailable,
                                        "HIP could not enumerate available devices");
  return static_cast<std::int32_t>(count);
}

HipMemoryResource::HipMemoryResource(std::int32_t device, MemorySpace memory_space) noexcept
    : device_(device), memory_space_(memory_space) {}
HipMemoryResource::~HipMemoryResource() = default;

Result<std::unique_ptr<HipMemoryResource>> HipMemoryResource::Create(std::int32_t device,
                                                                     MemorySpace memory_space) {
  if (memory_space != MemorySpace::kPinnedHost && memory_space != MemorySpace::kDevice &&
      memory_space != MemorySpace::kManaged)
    return Status(ErrorCode::kInvalidArgument,
                  "HipMemoryResource requires pinned-host, device, or managed memory");
  if (Status s = internal_core_hip::ValidateDevice(device); !s.ok()) return s;
  return std::unique_ptr<HipMemoryResource>(new HipMemoryResource(device, memory_space));
}

Result<void*> HipMemoryResource::Allocate(std::size_t bytes, std::size_t alignment) {
  if (alignment == 0 || (alignment & (alignment - 1)) != 0)
    return Status(ErrorCode::kInvalidArgument, "Allocation alignment must be a nonzero power of two");
  if (bytes == 0) return static_cast<void*>(nullptr);
  internal_core_hip::DeviceGuard guard(device_);
  if (Status s = guard.status(); !s.ok()) return s;
  void* p = nullptr;
  hipError_t e = hipSuccess;
  switch (memory_space_) {
    case MemorySpace::kPinnedHost: e = hipHostMalloc(&p, bytes, hipHostMallocDefault); break;
    case MemorySpace::kManaged: e = hipMallocManaged(&p, bytes, hipMemAttachGlobal); break;
    default: e = hipMalloc(&p, bytes); break;
  }
  if (e != hipSuccess)
    return internal_core_hip::HipStatus(e, ErrorCode::kAllocation, "HIP memory allocation failed");
  if (reinterpret_cast<std::uintptr_t>(p) % alignment != 0) {
    Deallocate(p, bytes, alignment);
    return Status(ErrorCode::kAllocation, "HIP could not satisfy the requested allocation alignment",
                  "hip");
  }
  return p;
}

void HipMemoryResource::Deallocate(void* pointer, std::size_t, std::size_t) noexcept {
  if (pointer == nullptr) return;
  internal_core_hip::DeviceGuard guard(device_);
  if (!guard.status().ok()) return;
  if (memory_space_ == MemorySpace::kPinnedHost) (void)hipHostFree(pointer);
  else (void)hipFree(pointer);
}

Result<ExecutionContext> CreateHipExecutionContext(std::int32_t device, Determinism determinism) {
  if (determinism != Determinism::kDeterministic && determinism != Determinism::kBackendDefault)
    return Status(ErrorCode::kInvalidArgument, "Determinism is not a recognized enumerator");
  if (Status s = internal_core_hip::ValidateDevice(device); !s.ok()) return s;
  internal_core_hip::DeviceGuard guard(device);
  if (Status s = guard.status(); !s.ok()) return s;
  hipStream_t stream = nullptr;
  const hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  if (e != hipSuccess)
    return internal_core_hip::HipStatus(e, ErrorCode::kProvider,
                                        "HIP could not create a nonblocking execution stream");
  auto state = std::make_shared<internal_core_hip::HipExecutionState>(device, stream);
  return internal_core_execution::Access::MakeContext(Backend::kHip, Device{Backend::kHip, device},
                                                      determinism, std::move(state));
}

Result<CompletionEvent> RecordHipEvent(const ExecutionContext& context) {
  if (context.backend() != Backend::kHip)
    return Status(ErrorCode::kInvalidArgument, "RecordHipEvent requires a HIP execution context");
  const auto& state = internal_core_execution::Access::State(context);
  if (state == nullptr)
    return Status(ErrorCode::kUnavailable, "The HIP execution context has no provider state");
  auto event = state->RecordEvent();
  if (!event.ok()) return event.status();
  return internal_core_execution::Access::MakeEvent(std::move(*event));
}

}  // namespace asc
