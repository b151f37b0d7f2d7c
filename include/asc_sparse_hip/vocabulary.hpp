// The asc:: vocabulary the HIP sparse facet (asc_sparse_hip/hip.hpp) is
// written against.
//
// Inside the asc-cpp tree (ASC_SPARSE_HIP_IN_ASC_TREE defined, as the
// ASC::sparse_hip target does for itself and its consumers) this header only
// includes the reference's own public headers, so the facet uses the
// reference's Status, Result, MemoryResource, Buffer, ExecutionContext,
// CompletionEvent and sparse views, unchanged.
//
// Outside it, the same names are declared here from scratch with the shapes
// of the reference's public declarations (include/asc/core/status.h,
// result.h, memory.h, execution.h, types.h; include/asc/sparse/coordinate.h,
// compressed.h, blas.h): the same class kinds, constructors, member
// functions, enumerator values and the provider-access seams
// (internal_sparse_compressed::ProviderAccess, internal_sparse_coordinate::
// ProviderAccess, internal_sparse_standard_blas::ProviderAccess,
// internal_core_execution::Access) that the reference's CUDA facet uses, so
// facet code compiles unchanged in both modes.  Only what the facet and its
// tests use is declared; out-of-line reference definitions are restated
// inline.
#pragma once

#if defined(ASC_SPARSE_HIP_IN_ASC_TREE)

#include "asc/core/execution.h"
#include "asc/core/memory.h"
#include "asc/core/result.h"
#include "asc/core/status.h"
#include "asc/core/types.h"
#include "asc/sparse/blas.h"
#include "asc/sparse/compressed.h"
#include "asc/sparse/coordinate.h"

#else  // standalone declarations

#include <algorithm>
#include <array>
#include <complex>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <optional>
#include <span>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>

namespace asc {

// ---- include/asc/core/types.h -------------------------------------------------
using extent_t = std::int64_t;
using stride_t = std::int64_t;
using index_t = std::int64_t;
using nnz_t = std::int64_t;
using rank_t = std::size_t;

// ---- include/asc/core/status.h:35-55 ----------------------------------------
enum class ErrorCode : std::uint32_t {
  kOk = 0,
  kInvalidArgument = 1,
  kShape = 2,
  kIndex = 3,
  kOverflow = 4,
  kInvalidState = 5,
  kAllocation = 6,
  kMemoryAccess = 7,
  kMemoryTransfer = 8,
  kUnsupported = 9,
  kUnavailable = 10,
  kProvider = 11,
  kNumerical = 12,
  kConfiguration = 13,
  kIo = 14,
  kEndOfFile = 15,
  kEncoding = 16,
  kVersion = 17,
  kInternal = 18,
};

class [[nodiscard]] Status {
 public:
  Status() noexcept = default;
  explicit Status(ErrorCode code, std::string message = {}, std::string provider = {},
                  std::int64_t native_code = 0) noexcept
      : code_(code), message_(std::move(message)), provider_(std::move(provider)),
        native_code_(native_code) {}
  static Status Ok() noexcept { return Status(); }
  [[nodiscard]] bool ok() const noexcept { return code_ == ErrorCode::kOk; }
  [[nodiscard]] ErrorCode code() const noexcept { return code_; }
  [[nodiscard]] const std::string& message() const noexcept { return message_; }
  [[nodiscard]] const std::string& provider() const noexcept { return provider_; }
  [[nodiscard]] std::int64_t native_code() const noexcept { return native_code_; }

 private:
  ErrorCode code_ = ErrorCode::kOk;
  std::string message_;
  std::string provider_;
  std::int64_t native_code_ = 0;
};

// ---- include/asc/core/result.h: one T or one non-OK Status ------------------
template <typename T>
class [[nodiscard]] Result {
  static_assert(!std::is_reference_v<T> && !std::is_void_v<T>);

 public:
  // Implicit, as the reference's, so functions can return either.
  Result(const T& value)  // NOLINT(google-explicit-constructor)
    requires std::copy_constructible<T>
      : value_(value) {}
  Result(T&& value) noexcept(std::is_nothrow_move_constructible_v<T>)  // NOLINT
      : value_(std::move(value)) {}
  Result(Status status)  // NOLINT(google-explicit-constructor)
    requires(!std::same_as<T, Status>)
      : status_(std::move(status)) {
    if (status_.ok()) Abort("a failed Result requires a non-OK Status");
  }
  static Result Failure(Status status) { return Result(std::move(status)); }
  [[nodiscard]] bool ok() const noexcept { return value_.has_value(); }
  [[nodiscard]] const Status& status() const noexcept { return status_; }
  [[nodiscard]] T& value() & { return Checked(); }
  [[nodiscard]] const T& value() const& { return Checked(); }
  [[nodiscard]] T&& value() && { return std::move(Checked()); }
  T& operator*() & { return Checked(); }
  const T& operator*() const& { return Checked(); }
  T&& operator*() && { return std::move(Checked()); }
  T* operator->() { return &Checked(); }
  const T* operator->() const { return &Checked(); }

 private:
  [[noreturn]] static void Abort(const char* what) {
    std::fprintf(stderr, "asc::Result contract violation: %s\n", what);
    std::abort();
  }
  T& Checked() {
    if (!value_) Abort(status_.message().c_str());
    return *value_;
  }
  const T& Checked() const {
    if (!value_) Abort(status_.message().c_str());
    return *value_;
  }
  std::optional<T> value_;
  Status status_;
};

// ---- include/asc/core/memory.h ----------------------------------------------
enum class MemorySpace : std::uint8_t {
  kHost = 0,
  kPinnedHost = 1,
  kDevice = 2,
  kManaged = 3,
};

class ConstMemoryView {
 public:
  constexpr ConstMemoryView(const void* data, std::size_t size, MemorySpace space) noexcept
      : data_(data), size_(size), space_(space) {}
  [[nodiscard]] constexpr const void* data() const noexcept { return data_; }
  [[nodiscard]] constexpr std::size_t size() const noexcept { return size_; }
  [[nodiscard]] constexpr MemorySpace space() const noexcept { return space_; }
  [[nodiscard]] constexpr bool valid() const noexcept { return size_ == 0 || data_ != nullptr; }

 private:
  const void* data_;
  std::size_t size_;
  MemorySpace space_;
};

class MutableMemoryView {
 public:
  constexpr MutableMemoryView(void* data, std::size_t size, MemorySpace space) noexcept
      : data_(data), size_(size), space_(space) {}
  [[nodiscard]] constexpr void* data() const noexcept { return data_; }
  [[nodiscard]] constexpr std::size_t size() const noexcept { return size_; }
  [[nodiscard]] constexpr MemorySpace space() const noexcept { return space_; }
  [[nodiscard]] constexpr bool valid() const noexcept { return size_ == 0 || data_ != nullptr; }
  constexpr operator ConstMemoryView() const noexcept {  // NOLINT(google-explicit-constructor)
    return {data_, size_, space_};
  }

 private:
  void* data_;
  std::size_t size_;
  MemorySpace space_;
};

// A resource must outlive every Buffer allocated from it.
class MemoryResource {
 public:
  virtual ~MemoryResource() = default;
  [[nodiscard]] virtual MemorySpace space() const noexcept = 0;
  virtual Result<void*> Allocate(std::size_t bytes, std::size_t alignment) = 0;
  virtual void Deallocate(void* pointer, std::size_t bytes, std::size_t alignment) noexcept = 0;
};

// One allocation of a MemoryResource, released once by its owner.
class Buffer {
 public:
  static Result<Buffer> Allocate(MemoryResource& resource, std::size_t bytes,
                                 std::size_t alignment = alignof(std::max_align_t)) {
    if (alignment == 0 || (alignment & (alignment - 1)) != 0)
      return Status(ErrorCode::kInvalidArgument, "Buffer alignment must be a nonzero power of two");
    auto allocation = resource.Allocate(bytes, alignment);
    if (!allocation.ok()) return allocation.status();
    void* p = *allocation;
    if (bytes != 0 && p == nullptr)
      return Status(ErrorCode::kAllocation, "MemoryResource returned null for a nonzero allocation");
    if (bytes == 0 && p != nullptr) {
      resource.Deallocate(p, bytes, alignment);
      return Status(ErrorCode::kInternal, "MemoryResource allocated storage for a zero-byte request");
    }
    if (p != nullptr && reinterpret_cast<std::uintptr_t>(p) % alignment != 0) {
      resource.Deallocate(p, bytes, alignment);
      return Status(ErrorCode::kAllocation, "MemoryResource returned a misaligned allocation");
    }
    return Buffer(&resource, p, bytes, alignment);
  }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  Buffer(Buffer&& other) noexcept
      : resource_(std::exchange(other.resource_, nullptr)),
        data_(std::exchange(other.data_, nullptr)),
        size_(std::exchange(other.size_, 0)),
        alignment_(std::exchange(other.alignment_, 0)) {}
  Buffer& operator=(Buffer&& other) noexcept {
    if (this != &other) {
      Reset();
      resource_ = std::exchange(other.resource_, nullptr);
      data_ = std::exchange(other.data_, nullptr);
      size_ = std::exchange(other.size_, 0);
      alignment_ = std::exchange(other.alignment_, 0);
    }
    return *this;
  }
  ~Buffer() { Reset(); }
  [[nodiscard]] bool valid() const noexcept { return resource_ != nullptr; }
  [[nodiscard]] void* data() noexcept { return data_; }
  [[nodiscard]] const void* data() const noexcept { return data_; }
  [[nodiscard]] std::size_t size() const noexcept { return size_; }
  [[nodiscard]] std::size_t alignment() const noexcept { return alignment_; }
  [[nodiscard]] Result<MemorySpace> space() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "A moved-from Buffer has no MemoryResource");
    return resource_->space();
  }
  [[nodiscard]] Result<MutableMemoryView> mutable_view() {
    auto s = space();
    if (!s.ok()) return s.status();
    return MutableMemoryView(data_, size_, *s);
  }
  [[nodiscard]] Result<ConstMemoryView> const_view() const {
    auto s = space();
    if (!s.ok()) return s.status();
    return ConstMemoryView(data_, size_, *s);
  }
  void Reset() noexcept {
    if (resource_ != nullptr && data_ != nullptr) resource_->Deallocate(data_, size_, alignment_);
    resource_ = nullptr;
    data_ = nullptr;
    size_ = 0;
    alignment_ = 0;
  }

 private:
  Buffer(MemoryResource* resource, void* data, std::size_t size, std::size_t alignment) noexcept
      : resource_(resource), data_(data), size_(size), alignment_(alignment) {}
  MemoryResource* resource_ = nullptr;
  void* data_ = nullptr;
  std::size_t size_ = 0;
  std::size_t alignment_ = 0;
};

// ---- include/asc/core/execution.h -------------------------------------------
enum class Backend : std::uint8_t { kSerial = 0, kCuda = 1, kHip = 2, kSycl = 3 };

struct Device {
  Backend backend = Backend::kSerial;
  std::int32_t ordinal = 0;
  static Device Serial() noexcept { return Device{Backend::kSerial, 0}; }
  friend bool operator==(const Device&, const Device&) = default;
};

enum class Determinism : std::uint8_t { kDeterministic = 0, kBackendDefault = 1 };

class ExecutionContext;
class CompletionEvent;

// The provider seam of src/core/execution_internal.h: a backend's execution
// state behind an ExecutionContext and its completion tokens.
namespace internal_core_execution {

class CompletionState {
 public:
  CompletionState() = default;
  CompletionState(const CompletionState&) = delete;
  CompletionState& operator=(const CompletionState&) = delete;
  virtual ~CompletionState() = default;
  [[nodiscard]] virtual Result<bool> Query() const = 0;
  [[nodiscard]] virtual Status Wait() const = 0;
};

class ExecutionState {
 public:
  ExecutionState() = default;
  ExecutionState(const ExecutionState&) = delete;
  ExecutionState& operator=(const ExecutionState&) = delete;
  virtual ~ExecutionState() = default;
  [[nodiscard]] virtual bool CanAccess(MemorySpace space) const noexcept = 0;
  [[nodiscard]] virtual Result<std::unique_ptr<CompletionState>> CopyBytes(
      MutableMemoryView destination, ConstMemoryView source, std::size_t byte_count) const = 0;
  [[nodiscard]] virtual Result<std::unique_ptr<CompletionState>> RecordEvent() const = 0;
  [[nodiscard]] virtual void* NativeExecutionHandle(Backend backend) const noexcept = 0;
};

class Access;

}  // namespace internal_core_execution

Result<CompletionEvent> CopyBytes(const ExecutionContext& context, MutableMemoryView destination,
                                  ConstMemoryView source, std::size_t byte_count);

class ExecutionContext {
 public:
  static ExecutionContext Serial() noexcept {
    return {Backend::kSerial, Device::Serial(), Determinism::kDeterministic};
  }
  // Serial only: other backends come from their provider's factory
  // (CreateHipExecutionContext for HIP), as in the reference.
  static Result<ExecutionContext> Create(Backend backend, Device device,
                                         Determinism determinism = Determinism::kDeterministic) {
    if (device.backend != backend)
      return Status(ErrorCode::kInvalidArgument, "Device backend does not match requested backend");
    if (device.ordinal < 0)
      return Status(ErrorCode::kInvalidArgument, "Device ordinal cannot be negative");
    if (backend == Backend::kSerial) {
      if (device.ordinal != 0)
        return Status(ErrorCode::kUnavailable, "Only serial CPU device ordinal zero is available");
      return ExecutionContext(backend, device, determinism);
    }
    return Status(ErrorCode::kUnsupported, "The requested backend has no approved core provider");
  }
  [[nodiscard]] Backend backend() const noexcept { return backend_; }
  [[nodiscard]] Device device() const noexcept { return device_; }
  [[nodiscard]] Determinism determinism() const noexcept { return determinism_; }
  [[nodiscard]] bool CanAccess(MemorySpace space) const noexcept {
    if (backend_ == Backend::kSerial) return space == MemorySpace::kHost;
    return state_ != nullptr && state_->CanAccess(space);
  }

 private:
  friend class internal_core_execution::Access;
  friend Result<CompletionEvent> CopyBytes(const ExecutionContext&, MutableMemoryView,
                                           ConstMemoryView, std::size_t);
  ExecutionContext(Backend backend, Device device, Determinism determinism,
                   std::shared_ptr<const internal_core_execution::ExecutionState> state =
                       nullptr) noexcept
      : backend_(backend), device_(device), determinism_(determinism), state_(std::move(state)) {}
  Backend backend_;
  Device device_;
  Determinism determinism_;
  std::shared_ptr<const internal_core_execution::ExecutionState> state_;
};

class CompletionEvent {
 public:
  CompletionEvent(const CompletionEvent&) = delete;
  CompletionEvent& operator=(const CompletionEvent&) = delete;
  CompletionEvent(CompletionEvent&& other) noexcept
      : valid_(std::exchange(other.valid_, false)),
        complete_(std::exchange(other.complete_, false)),
        state_(std::move(other.state_)) {}
  CompletionEvent& operator=(CompletionEvent&& other) noexcept {
    if (this != &other) {
      valid_ = std::exchange(other.valid_, false);
      complete_ = std::exchange(other.complete_, false);
      state_ = std::move(other.state_);
    }
    return *this;
  }
  ~CompletionEvent() = default;
  [[nodiscard]] bool valid() const noexcept { return valid_; }
  [[nodiscard]] bool is_complete() const noexcept { return valid_ && complete_; }
  [[nodiscard]] Result<bool> Query() const {
    if (!valid_)
      return Status(ErrorCode::kInvalidState, "A moved-from CompletionEvent cannot be queried");
    if (complete_) return true;
    if (state_ == nullptr)
      return Status(ErrorCode::kInternal, "A pending CompletionEvent has no provider state");
    auto q = state_->Query();
    if (!q.ok()) return q.status();
    complete_ = *q;
    return *q;
  }
  [[nodiscard]] Status Wait() const {
    if (!valid_)
      return Status(ErrorCode::kInvalidState, "A moved-from CompletionEvent cannot be waited on");
    if (complete_) return Status::Ok();
    if (state_ == nullptr)
      return Status(ErrorCode::kInternal, "A pending CompletionEvent has no provider state");
    Status s = state_->Wait();
    if (!s.ok()) return s;
    complete_ = true;
    return Status::Ok();
  }

 private:
  friend class internal_core_execution::Access;
  friend Result<CompletionEvent> CopyBytes(const ExecutionContext&, MutableMemoryView,
                                           ConstMemoryView, std::size_t);
  explicit CompletionEvent(bool complete) noexcept : valid_(true), complete_(complete) {}
  explicit CompletionEvent(std::unique_ptr<internal_core_execution::CompletionState> state) noexcept
      : valid_(state != nullptr), state_(std::move(state)) {}
  bool valid_ = false;
  mutable bool complete_ = false;
  std::unique_ptr<internal_core_execution::CompletionState> state_;
};

namespace internal_core_execution {

class Access {
 public:
  static ExecutionContext MakeContext(Backend backend, Device device, Determinism determinism,
                                      std::shared_ptr<const ExecutionState> state) noexcept {
    return {backend, device, determinism, std::move(state)};
  }
  static CompletionEvent MakeEvent(std::unique_ptr<CompletionState> state) noexcept {
    return CompletionEvent(std::move(state));
  }
  static CompletionEvent MakeCompletedEvent() noexcept { return CompletionEvent(true); }
  [[nodiscard]] static const std::shared_ptr<const ExecutionState>& State(
      const ExecutionContext& context) noexcept {
    return context.state_;
  }
};

}  // namespace internal_core_execution

// Stream-ordered byte copy through the context's provider (serial: memmove).
inline Result<CompletionEvent> CopyBytes(const ExecutionContext& context,
                                         MutableMemoryView destination, ConstMemoryView source,
                                         std::size_t byte_count) {
  if (byte_count > source.size() || byte_count > destination.size())
    return Status(ErrorCode::kMemoryAccess, "Copy byte count exceeds a memory view");
  if (!source.valid() || !destination.valid())
    return Status(ErrorCode::kMemoryAccess, "A nonempty memory view has a null address");
  if (!context.CanAccess(source.space()) || !context.CanAccess(destination.space())) {
    if (context.backend() == Backend::kSerial)
      return Status(ErrorCode::kUnsupported, "Serial execution can copy only host memory");
    return Status(ErrorCode::kMemoryAccess, "The execution context cannot access a memory view");
  }
  if (byte_count == 0) return CompletionEvent(true);
  if (context.backend() == Backend::kSerial) {
    if (source.data() != destination.data()) std::memmove(destination.data(), source.data(), byte_count);
    return CompletionEvent(true);
  }
  if (context.state_ == nullptr)
    return Status(ErrorCode::kUnavailable, "The execution context has no provider state");
  auto state = context.state_->CopyBytes(destination, source, byte_count);
  if (!state.ok()) return state.status();
  return CompletionEvent(std::move(*state));
}
inline Result<CompletionEvent> CopyBytes(const ExecutionContext& context,
                                         MutableMemoryView destination, ConstMemoryView source) {
  return CopyBytes(context, destination, source, source.size());
}

// ---- include/asc/sparse/coordinate.h ----------------------------------------
template <typename T>
concept SparseElement = (std::is_arithmetic_v<T> || std::same_as<T, std::complex<float>> ||
                         std::same_as<T, std::complex<double>>) &&
                        !std::same_as<T, bool>;

enum class DuplicatePolicy : std::uint8_t { kReject, kSum };
enum class ExplicitZeroPolicy : std::uint8_t { kKeep, kDrop };
enum class SparseCompressedFormat : std::uint8_t { kCsr, kCsc };

namespace internal_hip_vocabulary {

template <typename T>
concept ViewElement = SparseElement<std::remove_const_t<T>> && !std::is_volatile_v<T>;

inline bool SpansOverlap(const void* a, std::size_t an, const void* b, std::size_t bn) noexcept {
  if (an == 0 || bn == 0) return false;
  const auto x = reinterpret_cast<std::uintptr_t>(a), y = reinterpret_cast<std::uintptr_t>(b);
  if (an > UINTPTR_MAX - x || bn > UINTPTR_MAX - y) return true;
  return x < y + bn && y < x + an;
}

inline Status CheckSpan(const void* p, std::size_t bytes) {
  if (bytes == 0) return Status();
  if (p == nullptr) return Status(ErrorCode::kInvalidArgument, "Nonempty sparse storage cannot be null");
  if (bytes > UINTPTR_MAX - reinterpret_cast<std::uintptr_t>(p))
    return Status(ErrorCode::kOverflow, "Sparse storage address range overflows uintptr_t");
  return Status();
}

inline Result<std::size_t> Bytes(nnz_t count, std::size_t each) {
  if (count < 0) return Status(ErrorCode::kShape, "Sparse counts must be nonnegative");
  if (static_cast<std::uint64_t>(count) > SIZE_MAX / each)
    return Status(ErrorCode::kOverflow, "Sparse byte count overflows");
  return static_cast<std::size_t>(count) * each;
}

// Product of nonnegative extents, or the Status that rules them out.
template <std::size_t Rank>
inline Result<nnz_t> LogicalSize(std::span<const extent_t, Rank> extents) {
  nnz_t n = 1;
  for (extent_t e : extents) {
    if (e < 0) return Status(ErrorCode::kShape, "Sparse extents must be nonnegative");
    if (e != 0 && n > INT64_MAX / e) return Status(ErrorCode::kOverflow, "Sparse logical size overflows");
    n *= e;
  }
  return n;
}

// Rebinding a view's values (compressed.h:472-516, coordinate.h:533-569): a
// non-wrapping span, disjoint from the structure and not partially
// overlapping the current values.
inline Status CheckRebind(const void* values, const void* current, std::size_t value_bytes,
                          const void* first, std::size_t first_bytes, const void* second,
                          std::size_t second_bytes, const char* structure_message) {
  Status s = CheckSpan(values, value_bytes);
  if (!s.ok()) return s;
  if (SpansOverlap(first, first_bytes, values, value_bytes) ||
      SpansOverlap(second, second_bytes, values, value_bytes))
    return Status(ErrorCode::kInvalidArgument, structure_message);
  if (values != current && SpansOverlap(current, value_bytes, values, value_bytes))
    return Status(ErrorCode::kInvalidArgument,
                  "Rebound sparse values cannot partially overlap existing values");
  return Status();
}

}  // namespace internal_hip_vocabulary

namespace internal_sparse_coordinate {
class ProviderAccess;
}

// Canonical coordinate view: Rank indices per stored entry, interleaved, in
// strictly increasing lexicographic order.  Host views are checked by
// Create; other placements are trusted only when a provider made them.
template <internal_hip_vocabulary::ViewElement Element, std::size_t Rank>
class CoordinateView {
 public:
  using element_type = Element;
  using value_type = std::remove_const_t<Element>;
  static constexpr std::size_t kRank = Rank;

  static Result<CoordinateView> Create(const index_t* coordinates, Element* values,
                                       std::span<const extent_t, Rank> extents, nnz_t nonzeros,
                                       MemorySpace memory_space) {
    auto logical = internal_hip_vocabulary::LogicalSize<Rank>(extents);
    if (!logical.ok()) return logical.status();
    if (nonzeros < 0 || nonzeros > *logical)
      return Status(ErrorCode::kShape, "Sparse NNZ is incompatible with coordinate extents");
    auto cb = internal_hip_vocabulary::Bytes(nonzeros, Rank * sizeof(index_t));
    auto vb = internal_hip_vocabulary::Bytes(nonzeros, sizeof(value_type));
    if (!cb.ok()) return cb.status();
    if (!vb.ok()) return vb.status();
    for (const auto& [p, n] : {std::pair<const void*, std::size_t>{coordinates, *cb}, {values, *vb}}) {
      Status s = internal_hip_vocabulary::CheckSpan(p, n);
      if (!s.ok()) return s;
    }
    if (internal_hip_vocabulary::SpansOverlap(coordinates, *cb, values, *vb))
      return Status(ErrorCode::kInvalidArgument, "Coordinate storage spans cannot overlap");
    std::array<extent_t, Rank> copied{};
    std::copy(extents.begin(), extents.end(), copied.begin());
    CoordinateView v(coordinates, values, copied, nonzeros, memory_space, false);
    if (memory_space == MemorySpace::kHost) {
      for (nnz_t k = 0; k < nonzeros; ++k) {
        const index_t* c = coordinates + k * static_cast<nnz_t>(Rank);
        for (std::size_t d = 0; d < Rank; ++d)
          if (c[d] < 0 || c[d] >= copied[d])
            return Status(ErrorCode::kIndex, "Sparse coordinate is out of range");
        if (k > 0 && !std::lexicographical_compare(c - Rank, c, c, c + Rank))
          return Status(ErrorCode::kInvalidArgument,
                        "Sparse coordinates must be strictly increasing (canonical order)");
      }
      v.canonical_structure_trusted_ = true;
    }
    return v;
  }

  template <typename OtherElement>
    requires(std::is_const_v<Element> && std::same_as<std::remove_const_t<OtherElement>, value_type> &&
             !std::is_const_v<OtherElement>)
  constexpr CoordinateView(const CoordinateView<OtherElement, Rank>& other) noexcept  // NOLINT
      : coordinates_(other.coordinates()), values_(other.values()), extents_(other.shape()),
        nonzeros_(other.nnz()), memory_space_(other.memory_space()),
        canonical_structure_trusted_(other.canonical_structure_trusted()) {}

  [[nodiscard]] constexpr std::span<const extent_t, Rank> extents() const noexcept { return extents_; }
  [[nodiscard]] constexpr const std::array<extent_t, Rank>& shape() const noexcept { return extents_; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr const index_t* coordinates() const noexcept { return coordinates_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }
  [[nodiscard]] constexpr bool canonical_structure_trusted() const noexcept {
    return canonical_structure_trusted_;
  }

  template <internal_hip_vocabulary::ViewElement ReboundElement>
    requires std::same_as<std::remove_const_t<ReboundElement>, value_type>
  [[nodiscard]] Result<CoordinateView<ReboundElement, Rank>> RebindValues(
      ReboundElement* values) const {
    auto vb = internal_hip_vocabulary::Bytes(nonzeros_, sizeof(value_type));
    auto cb = internal_hip_vocabulary::Bytes(nonzeros_, Rank * sizeof(index_t));
    if (!vb.ok()) return vb.status();
    if (!cb.ok()) return cb.status();
    Status s = internal_hip_vocabulary::CheckRebind(values, values_, *vb, coordinates_, *cb,
                                                    nullptr, 0,
                                                    "Rebound sparse values cannot overlap coordinates");
    if (!s.ok()) return s;
    return CoordinateView<ReboundElement, Rank>(coordinates_, values, extents_, nonzeros_,
                                                memory_space_, canonical_structure_trusted_);
  }

 private:
  template <internal_hip_vocabulary::ViewElement, std::size_t>
  friend class CoordinateView;
  friend class internal_sparse_coordinate::ProviderAccess;
  constexpr CoordinateView(const index_t* coordinates, Element* values,
                           std::array<extent_t, Rank> extents, nnz_t nonzeros,
                           MemorySpace memory_space, bool trusted) noexcept
      : coordinates_(coordinates), values_(values), extents_(extents), nonzeros_(nonzeros),
        memory_space_(memory_space), canonical_structure_trusted_(trusted) {}
  const index_t* coordinates_;
  Element* values_;
  std::array<extent_t, Rank> extents_;
  nnz_t nonzeros_;
  MemorySpace memory_space_;
  bool canonical_structure_trusted_;
};

namespace internal_sparse_coordinate {
class ProviderAccess {
 public:
  template <internal_hip_vocabulary::ViewElement Element, std::size_t Rank>
  static CoordinateView<Element, Rank> MakeCanonicalView(const index_t* coordinates, Element* values,
                                                         std::array<extent_t, Rank> extents,
                                                         nnz_t nonzeros,
                                                         MemorySpace memory_space) noexcept {
    return CoordinateView<Element, Rank>(coordinates, values, extents, nonzeros, memory_space, true);
  }
};
}  // namespace internal_sparse_coordinate

// ---- include/asc/sparse/compressed.h ----------------------------------------
namespace internal_sparse_compressed {
class ProviderAccess;
}

// Canonical CSR/CSC view: outer offsets (outer+1), strictly increasing inner
// indices per segment, values.  Host views are checked by Create; other
// placements are trusted only when a provider made them.
template <internal_hip_vocabulary::ViewElement Element, SparseCompressedFormat Format>
class CompressedSparseView {
 public:
  using element_type = Element;
  using value_type = std::remove_const_t<Element>;
  static constexpr SparseCompressedFormat kFormat = Format;
  static constexpr std::size_t kRank = 2;

  static Result<CompressedSparseView> Create(const nnz_t* outer_offsets,
                                             const index_t* inner_indices, Element* values,
                                             std::span<const extent_t, 2> shape, nnz_t nonzeros,
                                             MemorySpace memory_space) {
    auto logical = internal_hip_vocabulary::LogicalSize<2>(shape);
    if (!logical.ok()) return logical.status();
    if (nonzeros < 0 || nonzeros > *logical)
      return Status(ErrorCode::kShape, "Sparse NNZ is incompatible with compressed shape");
    const extent_t outer = Format == SparseCompressedFormat::kCsr ? shape[0] : shape[1];
    const extent_t inner = Format == SparseCompressedFormat::kCsr ? shape[1] : shape[0];
    auto ob = internal_hip_vocabulary::Bytes(outer + 1, sizeof(nnz_t));
    auto ib = internal_hip_vocabulary::Bytes(nonzeros, sizeof(index_t));
    auto vb = internal_hip_vocabulary::Bytes(nonzeros, sizeof(value_type));
    if (!ob.ok()) return ob.status();
    if (!ib.ok()) return ib.status();
    if (!vb.ok()) return vb.status();
    if (outer_offsets == nullptr || (*ib != 0 && inner_indices == nullptr) ||
        (*vb != 0 && values == nullptr))
      return Status(ErrorCode::kInvalidArgument, "Compressed sparse storage pointers are incomplete");
    for (const auto& [p, n] : {std::pair<const void*, std::size_t>{outer_offsets, *ob},
                               {inner_indices, *ib}, {values, *vb}}) {
      Status s = internal_hip_vocabulary::CheckSpan(p, n);
      if (!s.ok()) return s;
    }
    using internal_hip_vocabulary::SpansOverlap;
    if (SpansOverlap(outer_offsets, *ob, inner_indices, *ib) ||
        SpansOverlap(outer_offsets, *ob, values, *vb) || SpansOverlap(inner_indices, *ib, values, *vb))
      return Status(ErrorCode::kInvalidArgument, "Compressed sparse storage spans cannot overlap");
    CompressedSparseView v(outer_offsets, inner_indices, values, {shape[0], shape[1]}, nonzeros,
                           memory_space, false);
    if (memory_space == MemorySpace::kHost) {
      if (outer_offsets[0] != 0 || outer_offsets[outer] != nonzeros)
        return Status(ErrorCode::kInvalidArgument,
                      "Compressed offsets must start at zero and end at NNZ");
      for (extent_t o = 0; o < outer; ++o) {
        if (outer_offsets[o + 1] < outer_offsets[o])
          return Status(ErrorCode::kInvalidArgument, "Compressed offsets must not decrease");
        for (nnz_t p = outer_offsets[o]; p < outer_offsets[o + 1]; ++p) {
          if (inner_indices[p] < 0 || inner_indices[p] >= inner)
            return Status(ErrorCode::kIndex, "Compressed inner index is out of range");
          if (p > outer_offsets[o] && inner_indices[p] <= inner_indices[p - 1])
            return Status(ErrorCode::kInvalidArgument,
                          "Compressed inner indices must strictly increase within a segment");
        }
      }
      v.canonical_structure_trusted_ = true;
    }
    return v;
  }

  template <typename OtherElement>
    requires(std::is_const_v<Element> && std::same_as<std::remove_const_t<OtherElement>, value_type> &&
             !std::is_const_v<OtherElement>)
  constexpr CompressedSparseView(const CompressedSparseView<OtherElement, Format>& other) noexcept  // NOLINT
      : outer_offsets_(other.outer_offsets()), inner_indices_(other.inner_indices()),
        values_(other.values()), shape_(other.shape()), nonzeros_(other.nnz()),
        memory_space_(other.memory_space()),
        canonical_structure_trusted_(other.canonical_structure_trusted()) {}

  [[nodiscard]] static constexpr rank_t rank() noexcept { return 2; }
  [[nodiscard]] static constexpr SparseCompressedFormat format() noexcept { return Format; }
  [[nodiscard]] constexpr const std::array<extent_t, 2>& shape() const noexcept { return shape_; }
  [[nodiscard]] constexpr std::span<const extent_t, 2> extents() const noexcept { return shape_; }
  [[nodiscard]] constexpr extent_t rows() const noexcept { return shape_[0]; }
  [[nodiscard]] constexpr extent_t columns() const noexcept { return shape_[1]; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr const nnz_t* outer_offsets() const noexcept { return outer_offsets_; }
  [[nodiscard]] constexpr const index_t* inner_indices() const noexcept { return inner_indices_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }
  [[nodiscard]] constexpr bool canonical_structure_trusted() const noexcept {
    return canonical_structure_trusted_;
  }

  template <internal_hip_vocabulary::ViewElement ReboundElement>
    requires std::same_as<std::remove_const_t<ReboundElement>, value_type>
  [[nodiscard]] Result<CompressedSparseView<ReboundElement, Format>> RebindValues(
      ReboundElement* values) const {
    const extent_t outer = Format == SparseCompressedFormat::kCsr ? shape_[0] : shape_[1];
    auto vb = internal_hip_vocabulary::Bytes(nonzeros_, sizeof(value_type));
    auto ob = internal_hip_vocabulary::Bytes(outer + 1, sizeof(nnz_t));
    auto ib = internal_hip_vocabulary::Bytes(nonzeros_, sizeof(index_t));
    if (!vb.ok()) return vb.status();
    if (!ob.ok()) return ob.status();
    if (!ib.ok()) return ib.status();
    Status s = internal_hip_vocabulary::CheckRebind(
        values, values_, *vb, outer_offsets_, *ob, inner_indices_, *ib,
        "Rebound sparse values cannot overlap compressed structure");
    if (!s.ok()) return s;
    return CompressedSparseView<ReboundElement, Format>(outer_offsets_, inner_indices_, values,
                                                        shape_, nonzeros_, memory_space_,
                                                        canonical_structure_trusted_);
  }

 private:
  template <internal_hip_vocabulary::ViewElement, SparseCompressedFormat>
  friend class CompressedSparseView;
  friend class internal_sparse_compressed::ProviderAccess;
  constexpr CompressedSparseView(const nnz_t* outer_offsets, const index_t* inner_indices,
                                 Element* values, std::array<extent_t, 2> shape, nnz_t nonzeros,
                                 MemorySpace memory_space, bool trusted) noexcept
      : outer_offsets_(outer_offsets), inner_indices_(inner_indices), values_(values),
        shape_(shape), nonzeros_(nonzeros), memory_space_(memory_space),
        canonical_structure_trusted_(trusted) {}
  const nnz_t* outer_offsets_;
  const index_t* inner_indices_;
  Element* values_;
  std::array<extent_t, 2> shape_;
  nnz_t nonzeros_;
  MemorySpace memory_space_;
  bool canonical_structure_trusted_;
};

namespace internal_sparse_compressed {
class ProviderAccess {
 public:
  template <internal_hip_vocabulary::ViewElement Element, SparseCompressedFormat Format>
  static CompressedSparseView<Element, Format> MakeCanonicalView(
      const nnz_t* outer_offsets, const index_t* inner_indices, Element* values,
      std::array<extent_t, 2> shape, nnz_t nonzeros, MemorySpace memory_space) noexcept {
    return CompressedSparseView<Element, Format>(outer_offsets, inner_indices, values, shape,
                                                 nonzeros, memory_space, true);
  }
};
}  // namespace internal_sparse_compressed

template <typename Element>
using CsrView = CompressedSparseView<Element, SparseCompressedFormat::kCsr>;
template <typename Element>
using CscView = CompressedSparseView<Element, SparseCompressedFormat::kCsc>;

// ---- include/asc/sparse/blas.h ----------------------------------------------
template <typename Element>
concept SparseBlasScalar = std::same_as<std::remove_cv_t<Element>, float> ||
                           std::same_as<std::remove_cv_t<Element>, double> ||
                           std::same_as<std::remove_cv_t<Element>, std::complex<float>> ||
                           std::same_as<std::remove_cv_t<Element>, std::complex<double>>;

enum class SparseBlasConjugation : std::uint8_t { kUnconjugated, kConjugated };
enum class SparseBlasTranspose : std::uint8_t { kNone, kTranspose, kConjugateTranspose };
enum class SparseBlasLayout : std::uint8_t { kColumnMajor, kRowMajor };
enum class SparseBlasTriangle : std::uint8_t { kUpper, kLower };
enum class SparseBlasDiagonal : std::uint8_t { kNonUnit, kUnit };

namespace internal_sparse_standard_blas {
class ProviderAccess;
}

// A strided vector over its backing storage: element i at
// logical_first + i * increment (a negative increment walks backwards); the
// reachable span must lie inside the backing storage.
template <typename Element>
class SparseBlasVectorView {
 public:
  using value_type = Element;
  using element_type = std::remove_const_t<Element>;
  static_assert(SparseBlasScalar<Element>, "SparseBlasVectorView requires a BLAS scalar");

  static Result<SparseBlasVectorView> Create(Element* logical_first, extent_t size,
                                             stride_t increment, ConstMemoryView backing_storage) {
    if (size < 0 || increment == 0 || !backing_storage.valid())
      return Status(ErrorCode::kInvalidArgument, "A sparse BLAS vector descriptor is invalid");
    if (size == 0) return SparseBlasVectorView(logical_first, 0, increment, backing_storage, nullptr, 0);
    const stride_t reach = (size - 1) * (increment < 0 ? -increment : increment);
    const Element* lo = increment < 0 ? logical_first - reach : logical_first;
    const std::size_t bytes = (static_cast<std::size_t>(reach) + 1) * sizeof(element_type);
    const auto b0 = reinterpret_cast<std::uintptr_t>(backing_storage.data());
    const auto l0 = reinterpret_cast<std::uintptr_t>(lo);
    if (logical_first == nullptr || l0 < b0 || l0 - b0 > backing_storage.size() ||
        bytes > backing_storage.size() - (l0 - b0))
      return Status(ErrorCode::kMemoryAccess, "A sparse BLAS vector exceeds its backing storage");
    return SparseBlasVectorView(logical_first, size, increment, backing_storage, lo, bytes);
  }
  operator SparseBlasVectorView<const element_type>() const noexcept  // NOLINT
    requires(!std::is_const_v<Element>)
  {
    return SparseBlasVectorView<const element_type>(data_, size_, increment_, backing_storage_,
                                                    reachable_data_, reachable_size_);
  }
  [[nodiscard]] Element* data() const noexcept { return data_; }
  [[nodiscard]] extent_t size() const noexcept { return size_; }
  [[nodiscard]] stride_t increment() const noexcept { return increment_; }
  [[nodiscard]] MemorySpace memory_space() const noexcept { return backing_storage_.space(); }
  [[nodiscard]] ConstMemoryView backing_storage() const noexcept { return backing_storage_; }
  [[nodiscard]] ConstMemoryView reachable_storage() const noexcept {
    return {reachable_data_, reachable_size_, backing_storage_.space()};
  }

 private:
  template <typename>
  friend class SparseBlasVectorView;
  SparseBlasVectorView(Element* data, extent_t size, stride_t increment, ConstMemoryView backing,
                       const void* reachable_data, std::size_t reachable_size) noexcept
      : data_(data), size_(size), increment_(increment), backing_storage_(backing),
        reachable_data_(reachable_data), reachable_size_(reachable_size) {}
  Element* data_;
  extent_t size_;
  stride_t increment_;
  ConstMemoryView backing_storage_;
  const void* reachable_data_;
  std::size_t reachable_size_;
};

// Canonical indexed vector: nnz strictly increasing indices in
// [0, dense_extent) and their values.  Host views are checked by Create.
template <typename Element>
class SparseBlasIndexedVectorView {
 public:
  using value_type = Element;
  using element_type = std::remove_const_t<Element>;
  static_assert(SparseBlasScalar<Element>, "SparseBlasIndexedVectorView requires a BLAS scalar");

  static Result<SparseBlasIndexedVectorView> Create(const index_t* indices, Element* values,
                                                    nnz_t nonzeros, extent_t dense_extent,
                                                    ConstMemoryView index_storage,
                                                    ConstMemoryView value_storage) {
    if (nonzeros < 0 || dense_extent < 0 || nonzeros > dense_extent ||
        index_storage.space() != value_storage.space())
      return Status(ErrorCode::kInvalidArgument, "A sparse indexed-vector descriptor is invalid");
    SparseBlasIndexedVectorView v(indices, values, nonzeros, dense_extent, index_storage,
                                  value_storage, false);
    if (index_storage.space() == MemorySpace::kHost) {
      index_t previous = -1;
      for (nnz_t p = 0; p < nonzeros; ++p) {
        if (indices[p] <= previous || indices[p] >= dense_extent)
          return Status(ErrorCode::kIndex,
                        "Sparse indexed-vector indices must be sorted, unique and in range");
        previous = indices[p];
      }
      v.canonical_structure_trusted_ = true;
    }
    return v;
  }
  operator SparseBlasIndexedVectorView<const element_type>() const noexcept  // NOLINT
    requires(!std::is_const_v<Element>)
  {
    return SparseBlasIndexedVectorView<const element_type>(indices_, values_, nonzeros_,
                                                           dense_extent_, index_storage_,
                                                           value_storage_,
                                                           canonical_structure_trusted_);
  }
  [[nodiscard]] const index_t* indices() const noexcept { return indices_; }
  [[nodiscard]] Element* values() const noexcept { return values_; }
  [[nodiscard]] nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] extent_t dense_extent() const noexcept { return dense_extent_; }
  [[nodiscard]] MemorySpace memory_space() const noexcept { return index_storage_.space(); }
  [[nodiscard]] ConstMemoryView index_storage() const noexcept { return index_storage_; }
  [[nodiscard]] ConstMemoryView value_storage() const noexcept { return value_storage_; }
  [[nodiscard]] bool canonical_structure_trusted() const noexcept {
    return canonical_structure_trusted_;
  }

 private:
  template <typename>
  friend class SparseBlasIndexedVectorView;
  friend class internal_sparse_standard_blas::ProviderAccess;
  SparseBlasIndexedVectorView(const index_t* indices, Element* values, nnz_t nonzeros,
                              extent_t dense_extent, ConstMemoryView index_storage,
                              ConstMemoryView value_storage, bool trusted) noexcept
      : indices_(indices), values_(values), nonzeros_(nonzeros), dense_extent_(dense_extent),
        index_storage_(index_storage), value_storage_(value_storage),
        canonical_structure_trusted_(trusted) {}
  const index_t* indices_;
  Element* values_;
  nnz_t nonzeros_;
  extent_t dense_extent_;
  ConstMemoryView index_storage_;
  ConstMemoryView value_storage_;
  bool canonical_structure_trusted_;
};

namespace internal_sparse_standard_blas {
class ProviderAccess {
 public:
  template <typename Element>
  static SparseBlasIndexedVectorView<Element> MakeCanonicalIndexedVector(
      const index_t* indices, Element* values, nnz_t nonzeros, extent_t dense_extent,
      ConstMemoryView index_storage, ConstMemoryView value_storage) noexcept {
    return SparseBlasIndexedVectorView<Element>(indices, values, nonzeros, dense_extent,
                                                index_storage, value_storage, true);
  }
};
}  // namespace internal_sparse_standard_blas

// Dense block, column- or row-major with a leading dimension, inside its
// backing storage.
template <typename Element>
class SparseBlasMatrixView {
 public:
  using value_type = Element;
  using element_type = std::remove_const_t<Element>;
  static_assert(SparseBlasScalar<Element>, "SparseBlasMatrixView requires a BLAS scalar");

  static Result<SparseBlasMatrixView> Create(Element* data, extent_t rows, extent_t columns,
                                             SparseBlasLayout layout, stride_t leading_dimension,
                                             ConstMemoryView backing_storage) {
    if (rows < 0 || columns < 0 ||
        (layout != SparseBlasLayout::kColumnMajor && layout != SparseBlasLayout::kRowMajor))
      return Status(ErrorCode::kInvalidArgument, "A sparse BLAS dense-matrix descriptor is invalid");
    const extent_t contiguous = layout == SparseBlasLayout::kColumnMajor ? rows : columns;
    if (leading_dimension < std::max<stride_t>(1, contiguous))
      return Status(ErrorCode::kInvalidArgument, "A sparse BLAS leading dimension is too small");
    std::size_t bytes = 0;
    if (rows != 0 && columns != 0) {
      const extent_t major = layout == SparseBlasLayout::kColumnMajor ? columns : rows;
      if (major - 1 > (INT64_MAX - contiguous) / leading_dimension)
        return Status(ErrorCode::kOverflow, "A sparse BLAS dense-matrix span overflows");
      const stride_t span = (major - 1) * leading_dimension + contiguous;
      if (static_cast<std::uint64_t>(span) > SIZE_MAX / sizeof(element_type))
        return Status(ErrorCode::kOverflow, "A sparse BLAS dense-matrix span overflows");
      bytes = static_cast<std::size_t>(span) * sizeof(element_type);
      const auto b0 = reinterpret_cast<std::uintptr_t>(backing_storage.data());
      const auto d0 = reinterpret_cast<std::uintptr_t>(data);
      if (data == nullptr || !backing_storage.valid() || d0 < b0 ||
          d0 - b0 > backing_storage.size() || bytes > backing_storage.size() - (d0 - b0))
        return Status(ErrorCode::kMemoryAccess, "A sparse BLAS dense matrix exceeds its backing storage");
    }
    return SparseBlasMatrixView(data, rows, columns, layout, leading_dimension, backing_storage,
                                bytes);
  }
  operator SparseBlasMatrixView<const element_type>() const noexcept  // NOLINT
    requires(!std::is_const_v<Element>)
  {
    return SparseBlasMatrixView<const element_type>(data_, rows_, columns_, layout_,
                                                    leading_dimension_, backing_storage_, bytes_);
  }
  [[nodiscard]] Element* data() const noexcept { return data_; }
  [[nodiscard]] extent_t rows() const noexcept { return rows_; }
  [[nodiscard]] extent_t columns() const noexcept { return columns_; }
  [[nodiscard]] SparseBlasLayout layout() const noexcept { return layout_; }
  [[nodiscard]] stride_t leading_dimension() const noexcept { return leading_dimension_; }
  [[nodiscard]] MemorySpace memory_space() const noexcept { return backing_storage_.space(); }
  [[nodiscard]] ConstMemoryView reachable_storage() const noexcept {
    return {data_, bytes_, backing_storage_.space()};
  }

 private:
  template <typename>
  friend class SparseBlasMatrixView;
  SparseBlasMatrixView(Element* data, extent_t rows, extent_t columns, SparseBlasLayout layout,
                       stride_t leading_dimension, ConstMemoryView backing, std::size_t bytes) noexcept
      : data_(data), rows_(rows), columns_(columns), layout_(layout),
        leading_dimension_(leading_dimension), backing_storage_(backing), bytes_(bytes) {}
  Element* data_;
  extent_t rows_;
  extent_t columns_;
  SparseBlasLayout layout_;
  stride_t leading_dimension_;
  ConstMemoryView backing_storage_;
  std::size_t bytes_;
};

}  // namespace asc

#endif  // ASC_SPARSE_HIP_IN_ASC_TREE
