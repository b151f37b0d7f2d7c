// ASC::core_hip: the HIP core provider behind the HIP sparse facet.
//
// The HIP counterpart of the reference's CUDA core provider
// (include/asc/core/providers/cuda.h, src/core/cuda/runtime.cc): a device
// MemoryResource, the factory of HIP ExecutionContexts and the event
// recorder.  A HIP ExecutionContext carries an execution state that
// implements the reference's ExecutionState seam (src/core/
// execution_internal.h:25-41: CanAccess, CopyBytes, RecordEvent,
// NativeExecutionHandle(kHip) = its hipStream_t), so the reference's
// CopyBytes and CompletionEvent work on it unchanged.  Defined in
// asc-cpp_amd/csrc/facet/core_hip.cc.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

#include "asc_sparse_hip/vocabulary.hpp"

namespace asc {

// Visible HIP devices (CudaDeviceCount, providers/cuda.h:30).
Result<std::int32_t> HipDeviceCount();

// Device, pinned-host or managed memory of one HIP device
// (CudaMemoryResource, providers/cuda.h:49-110).
class HipMemoryResource final : public MemoryResource {
 public:
  static Result<std::unique_ptr<HipMemoryResource>> Create(std::int32_t device,
                                                           MemorySpace memory_space);
  HipMemoryResource(const HipMemoryResource&) = delete;
  HipMemoryResource& operator=(const HipMemoryResource&) = delete;
  HipMemoryResource(HipMemoryResource&&) = delete;
  HipMemoryResource& operator=(HipMemoryResource&&) = delete;
  ~HipMemoryResource() override;

  [[nodiscard]] MemorySpace space() const noexcept override { return memory_space_; }
  [[nodiscard]] std::int32_t device() const noexcept { return device_; }
  Result<void*> Allocate(std::size_t bytes, std::size_t alignment) override;
  void Deallocate(void* pointer, std::size_t bytes, std::size_t alignment) noexcept override;

 private:
  HipMemoryResource(std::int32_t device, MemorySpace memory_space) noexcept;
  std::int32_t device_;
  MemorySpace memory_space_;
};

// A HIP execution context on `device` with its own non-blocking stream at
// default priority (CreateCudaExecutionContext, runtime.cc:434-465).  Where
// ExecutionContext::Create(Backend::kHip, ...) reports kUnsupported, this
// factory is the provider's entry point.
Result<ExecutionContext> CreateHipExecutionContext(
    std::int32_t device, Determinism determinism = Determinism::kDeterministic);

// Completion of everything enqueued so far on a HIP context's stream
// (RecordCudaEvent, runtime.cc:467-482).
Result<CompletionEvent> RecordHipEvent(const ExecutionContext& context);

}  // namespace asc
