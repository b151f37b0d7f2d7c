// SparseHipContext (include/asc_sparse_hip/hip.hpp): the HIP sparse
// provider's context over a HIP ExecutionContext, the counterpart of
// SparseCudaContext::Create (src/sparse/cuda/context.cc:69-113).  The
// library context (C ABI) borrows the execution context's stream, read
// through the reference's ExecutionState seam, so the facet's operations and
// the core's CopyBytes are ordered on one stream.
#include <utility>

#include "asc_sparse_hip/hip.hpp"
#if defined(ASC_SPARSE_HIP_IN_ASC_TREE)
#include "execution_internal.h"  // src/core/execution_internal.h of the asc-cpp tree
#endif

namespace asc {

Result<SparseHipContext> SparseHipContext::Create(ExecutionContext execution_context) {
  if (execution_context.backend() != Backend::kHip)
    return Status(ErrorCode::kInvalidArgument, "SparseHipContext requires a HIP execution context");
  const auto& state = internal_core_execution::Access::State(execution_context);
  if (state == nullptr)
    return Status(ErrorCode::kUnavailable, "The HIP execution context has no provider state");
  void* stream = state->NativeExecutionHandle(Backend::kHip);
  if (stream == nullptr)
    return Status(ErrorCode::kUnavailable, "The HIP execution context has no execution stream");
  const std::uint8_t determinism = execution_context.determinism() == Determinism::kDeterministic
                                       ? ASC_HIP_DETERMINISTIC
                                       : ASC_HIP_BACKEND_DEFAULT;
  asc_hip_context* raw = nullptr;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_context_create(
      execution_context.device().ordinal, stream, determinism, &raw));
  if (!s.ok()) return s;
  return SparseHipContext(std::move(execution_context), raw);
}

Result<SparseHipContext> SparseHipContext::CreateValidationOnly() {
  asc_hip_context* raw = nullptr;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_context_create(-1, nullptr, ASC_HIP_DETERMINISTIC, &raw));
  if (!s.ok()) return s;
  return SparseHipContext(ExecutionContext::Serial(), raw);
}

}  // namespace asc
