// ASC::sparse_hip: the C++20 facet of the MI355X Sparse BLAS provider.
//
// The public surface mirrors the reference's CUDA facet
// (include/asc/sparse/providers/cuda.h) with Cuda -> Hip: SparseHipContext
// (:499-606), HipStridedVectorView (:616-730), HipCsrArray / HipCsrClone /
// HipCloneCsr (:810-1056), HipCsrSpmvWorkspaceSize + HipCsrSpmv (:1758-1802),
// HipSpmv / HipSpmm (:1616-1658), each returning Result<...>.  It adds the
// device conversions the reference provider lacks (HipConvertToCsc,
// HipConvertToCsr, HipConvertToCoordinate) and the column-panel analysis.
//
// Every function packs plain descriptors and calls the C ABI of
// include/asc_sparse_hip.h, the type-erased seam that replaces the
// reference's *Erased functions (providers/cuda.h:212-253).  Nothing here
// throws; failures come back as Status with the asc::ErrorCode numbering,
// provider "hip" and the hipError_t as native code.  Enqueue success does not
// imply completion: wait on the returned CompletionEvent.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <type_traits>
#include <utility>

#include "asc_sparse_hip.h"
#include "asc_sparse_hip/vocabulary.hpp"

namespace asc {

// include/asc/core/execution.h:42-46
enum class Backend : std::uint8_t { kSerial = 0, kCuda = 1, kHip = 2 };

namespace internal_sparse_hip {

inline Status ToStatus(asc_hip_status s) {
  if (s.code == ASC_HIP_OK) return Status();
  return Status(static_cast<ErrorCode>(s.code), asc_sparse_hip_last_message(),
                s.provider == ASC_HIP_PROVIDER_HIP ? "hip" : "", s.native);
}

template <typename E>
inline constexpr bool kSupportedElement =
    std::is_same_v<std::remove_const_t<E>, float> || std::is_same_v<std::remove_const_t<E>, double>;

template <typename E>
constexpr std::uint8_t ElementKind() {
  return std::is_same_v<std::remove_const_t<E>, double> ? ASC_HIP_F64 : ASC_HIP_F32;
}

struct ContextDeleter {
  void operator()(asc_hip_context* c) const noexcept { (void)asc_sparse_hip_context_destroy(c); }
};
struct EventDeleter {
  void operator()(asc_hip_event* e) const noexcept { (void)asc_sparse_hip_event_destroy(e); }
};
using ContextHandle = std::shared_ptr<asc_hip_context>;

}  // namespace internal_sparse_hip

// Completion state of enqueued work (include/asc/core/execution.h:246-369):
// holds no arrays, only the event.
class CompletionEvent {
 public:
  CompletionEvent() = default;
  explicit CompletionEvent(std::shared_ptr<asc_hip_event> event) : event_(std::move(event)) {}
  [[nodiscard]] Status Wait() const {
    if (!event_) return Status();
    return internal_sparse_hip::ToStatus(asc_sparse_hip_event_wait(event_.get()));
  }
  [[nodiscard]] asc_hip_event* native() const noexcept { return event_.get(); }

 private:
  std::shared_ptr<asc_hip_event> event_;
};

// A device and the stream work is ordered on (the HIP counterpart of
// CreateCudaExecutionContext, src/core/cuda/runtime.cc:434-465).
class ExecutionContext {
 public:
  [[nodiscard]] Backend backend() const noexcept { return Backend::kHip; }
  [[nodiscard]] std::int32_t device() const noexcept { return device_; }
  // hipStream_t to borrow, or null for a stream owned by each provider context.
  [[nodiscard]] void* native_stream() const noexcept { return stream_; }

 private:
  friend Result<ExecutionContext> CreateHipExecutionContext(std::int32_t, void*);
  std::int32_t device_ = 0;
  void* stream_ = nullptr;
};

inline Result<std::int32_t> HipDeviceCount() { return asc_sparse_hip_device_count(); }

inline Result<ExecutionContext> CreateHipExecutionContext(std::int32_t device,
                                                          void* stream = nullptr) {
  const std::int32_t count = asc_sparse_hip_device_count();
  if (device < 0 || device >= count)
    return Status(ErrorCode::kUnavailable, "no such HIP device", "hip");
  ExecutionContext c;
  c.device_ = device;
  c.stream_ = stream;
  return c;
}

// Provider state bound to an execution context (SparseCudaContext,
// providers/cuda.h:499-606).  Move-only; one context must not be used from
// two threads at once.
class SparseHipContext {
 public:
  static Result<SparseHipContext> Create(ExecutionContext execution_context) {
    asc_hip_context* raw = nullptr;
    const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_context_create(
        execution_context.device(), execution_context.native_stream(), ASC_HIP_DETERMINISTIC,
        &raw));
    if (!s.ok()) return s;
    return SparseHipContext(execution_context,
                            internal_sparse_hip::ContextHandle(raw, internal_sparse_hip::ContextDeleter{}));
  }
  // A context that validates every call and enqueues nothing (no device is
  // touched): calls that pass validation return kUnavailable.
  static Result<SparseHipContext> CreateValidationOnly() {
    asc_hip_context* raw = nullptr;
    const Status s = internal_sparse_hip::ToStatus(
        asc_sparse_hip_context_create(-1, nullptr, ASC_HIP_DETERMINISTIC, &raw));
    if (!s.ok()) return s;
    return SparseHipContext(ExecutionContext(),
                            internal_sparse_hip::ContextHandle(raw, internal_sparse_hip::ContextDeleter{}));
  }
  SparseHipContext(const SparseHipContext&) = delete;
  SparseHipContext& operator=(const SparseHipContext&) = delete;
  SparseHipContext(SparseHipContext&&) noexcept = default;
  SparseHipContext& operator=(SparseHipContext&&) noexcept = default;

  [[nodiscard]] const ExecutionContext& execution_context() const noexcept { return execution_; }
  [[nodiscard]] asc_hip_context* native() const noexcept { return handle_.get(); }
  [[nodiscard]] const internal_sparse_hip::ContextHandle& handle() const noexcept { return handle_; }
  [[nodiscard]] Status Synchronize() const {
    return internal_sparse_hip::ToStatus(asc_sparse_hip_context_synchronize(handle_.get()));
  }
  // Records completion of everything enqueued so far.
  [[nodiscard]] Result<CompletionEvent> Record() const {
    asc_hip_event* e = nullptr;
    const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_event_record(handle_.get(), &e));
    if (!s.ok()) return s;
    return CompletionEvent(std::shared_ptr<asc_hip_event>(e, internal_sparse_hip::EventDeleter{}));
  }

 private:
  SparseHipContext(ExecutionContext e, internal_sparse_hip::ContextHandle h)
      : execution_(e), handle_(std::move(h)) {}
  ExecutionContext execution_;
  internal_sparse_hip::ContextHandle handle_;
};

// Device or pinned-host memory of one device (CudaMemoryResource,
// include/asc/core/providers/cuda.h:39-120).
class HipMemoryResource {
 public:
  static Result<HipMemoryResource> Create(const SparseHipContext& context, MemorySpace space) {
    if (space != MemorySpace::kDevice && space != MemorySpace::kPinnedHost)
      return Status(ErrorCode::kUnsupported, "HIP resources hold device or pinned host memory");
    return HipMemoryResource(context.handle(), space);
  }
  [[nodiscard]] Result<void*> Allocate(std::size_t bytes) const {
    void* p = nullptr;
    const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_allocate(
        handle_.get(), bytes, static_cast<std::uint8_t>(space_), &p));
    if (!s.ok()) return s;
    return p;
  }
  void Deallocate(void* p) const noexcept {
    if (p) (void)asc_sparse_hip_deallocate(handle_.get(), p, static_cast<std::uint8_t>(space_));
  }
  [[nodiscard]] MemorySpace memory_space() const noexcept { return space_; }

 private:
  HipMemoryResource(internal_sparse_hip::ContextHandle h, MemorySpace s)
      : handle_(std::move(h)), space_(s) {}
  internal_sparse_hip::ContextHandle handle_;
  MemorySpace space_;
};

// One owned allocation of a HipMemoryResource.
class HipBuffer {
 public:
  HipBuffer() = default;
  static Result<HipBuffer> Allocate(const HipMemoryResource& resource, std::size_t bytes) {
    auto p = resource.Allocate(bytes);
    if (!p.ok()) return p.status();
    HipBuffer b;
    b.ptr_ = std::shared_ptr<void>(*p, [resource](void* q) { resource.Deallocate(q); });
    b.size_ = bytes;
    b.space_ = resource.memory_space();
    return b;
  }
  [[nodiscard]] void* data() const noexcept { return ptr_.get(); }
  [[nodiscard]] std::size_t size() const noexcept { return size_; }
  [[nodiscard]] MutableMemoryView mutable_view() const noexcept { return {ptr_.get(), size_, space_}; }
  [[nodiscard]] ConstMemoryView const_view() const noexcept { return {ptr_.get(), size_, space_}; }

 private:
  std::shared_ptr<void> ptr_;
  std::size_t size_ = 0;
  MemorySpace space_ = MemorySpace::kDevice;
};

// Grants the facet the trusted-view constructor that the vocabulary keeps
// private (compressed.h:759-770).
class HipProviderAccess {
 public:
  template <typename E, SparseCompressedFormat F>
  static CompressedSparseView<E, F> TrustedDevice(const void* offsets, const void* indices,
                                                  E* values, extent_t rows, extent_t cols,
                                                  nnz_t nnz, int index_bytes) {
    CompressedSparseView<E, F> v;
    v.offsets_ = static_cast<const nnz_t*>(offsets);
    v.indices_ = static_cast<const index_t*>(indices);
    v.values_ = values;
    v.shape_ = {rows, cols};
    v.nonzeros_ = nnz;
    v.memory_space_ = MemorySpace::kDevice;
    v.trusted_ = true;
    v.index_bytes_ = index_bytes;
    return v;
  }
  template <typename E>
  static CoordinateView<E, 2> TrustedDeviceCoordinates(const void* coords, E* values, extent_t rows,
                                                       extent_t cols, nnz_t nnz) {
    CoordinateView<E, 2> v;
    v.coordinates_ = static_cast<const index_t*>(coords);
    v.values_ = values;
    v.extents_ = {rows, cols};
    v.nonzeros_ = nnz;
    v.memory_space_ = MemorySpace::kDevice;
    v.trusted_ = true;
    return v;
  }
  template <typename E, SparseCompressedFormat F>
  static asc_hip_sparse_desc Describe(const CompressedSparseView<E, F>& v) {
    asc_hip_sparse_desc d{};
    d.outer_offsets = v.outer_offsets();
    d.inner_indices = v.inner_indices();
    d.values = v.values();
    d.rows = v.rows();
    d.cols = v.columns();
    d.nnz = v.nnz();
    d.format = F == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC;
    d.element_kind = internal_sparse_hip::ElementKind<E>();
    d.index_bytes = static_cast<std::uint8_t>(v.index_bytes());
    d.offset_bytes = 8;
    d.memory_space = static_cast<std::uint8_t>(v.memory_space());
    d.trusted = v.canonical_structure_trusted() ? 1 : 0;
    return d;
  }
  template <typename E>
  static SparseBlasIndexedVectorView<E> TrustedDeviceIndexed(const void* indices, E* values,
                                                             nnz_t nnz, extent_t extent) {
    SparseBlasIndexedVectorView<E> v;
    v.indices_ = static_cast<const index_t*>(indices);
    v.values_ = values;
    v.nonzeros_ = nnz;
    v.dense_extent_ = extent;
    v.memory_space_ = MemorySpace::kDevice;
    v.trusted_ = true;
    return v;
  }
  template <typename E>
  static asc_hip_indexed_desc Describe(const SparseBlasIndexedVectorView<E>& v) {
    asc_hip_indexed_desc d{};
    d.indices = v.indices();
    d.values = v.values();
    d.nnz = v.nnz();
    d.dense_extent = v.dense_extent();
    d.element_kind = internal_sparse_hip::ElementKind<E>();
    d.memory_space = static_cast<std::uint8_t>(v.memory_space());
    d.writable = std::is_const_v<E> ? 0 : 1;
    d.trusted = v.canonical_structure_trusted() ? 1 : 0;
    return d;
  }
  template <typename E>
  static asc_hip_sparse_desc Describe(const CoordinateView<E, 2>& v) {
    asc_hip_sparse_desc d{};
    d.outer_offsets = v.coordinates();
    d.values = v.values();
    d.rows = v.extent(0);
    d.cols = v.extent(1);
    d.nnz = v.nnz();
    d.format = ASC_HIP_COO;
    d.element_kind = internal_sparse_hip::ElementKind<E>();
    d.index_bytes = 8;
    d.offset_bytes = 8;
    d.memory_space = static_cast<std::uint8_t>(v.memory_space());
    d.trusted = v.canonical_structure_trusted() ? 1 : 0;
    return d;
  }
};

// HIP-accessible strided vector (CudaStridedVectorView, providers/cuda.h:616-730).
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipStridedVectorView {
 public:
  static Result<HipStridedVectorView> Create(Element* data, extent_t extent, stride_t stride) {
    if (extent < 0 || stride <= 0)
      return Status(ErrorCode::kInvalidArgument,
                    "A HIP strided vector requires nonnegative extent and positive stride");
    HipStridedVectorView v;
    v.data_ = data;
    v.extent_ = extent;
    v.stride_ = stride;
    return v;
  }
  [[nodiscard]] constexpr Element* data() const noexcept { return data_; }
  [[nodiscard]] constexpr extent_t extent() const noexcept { return extent_; }
  [[nodiscard]] constexpr stride_t stride() const noexcept { return stride_; }
  [[nodiscard]] static constexpr MemorySpace memory_space() noexcept { return MemorySpace::kDevice; }
  operator HipStridedVectorView<const Element>() const noexcept {  // NOLINT
    return *HipStridedVectorView<const Element>::Create(data_, extent_, stride_);
  }

 private:
  Element* data_ = nullptr;
  extent_t extent_ = 0;
  stride_t stride_ = 1;
};

namespace internal_sparse_hip {

template <typename E>
asc_hip_vector_desc Describe(const HipStridedVectorView<E>& v) {
  asc_hip_vector_desc d{};
  d.data = v.data();
  d.extent = v.extent();
  d.stride = v.stride();
  d.element_kind = ElementKind<E>();
  d.memory_space = ASC_HIP_DEVICE;
  d.writable = std::is_const_v<E> ? 0 : 1;
  return d;
}
template <typename E>
asc_hip_vector_desc Describe(const SparseBlasVectorView<E>& v) {
  asc_hip_vector_desc d{};
  d.data = v.data();
  d.extent = v.size();
  d.stride = v.increment();
  d.element_kind = ElementKind<E>();
  d.memory_space = static_cast<std::uint8_t>(v.memory_space());
  d.writable = std::is_const_v<E> ? 0 : 1;
  return d;
}
template <typename E>
asc_hip_matrix_desc Describe(const SparseBlasMatrixView<E>& m) {
  asc_hip_matrix_desc d{};
  d.data = m.data();
  d.rows = m.rows();
  d.cols = m.columns();
  d.leading_dimension = m.leading_dimension();
  d.layout = m.layout() == SparseBlasLayout::kRowMajor ? ASC_HIP_ROW_MAJOR : ASC_HIP_COLUMN_MAJOR;
  d.element_kind = ElementKind<E>();
  d.memory_space = static_cast<std::uint8_t>(m.memory_space());
  d.writable = std::is_const_v<E> ? 0 : 1;
  return d;
}
inline std::uint8_t Op(SparseBlasTranspose t) {
  return t == SparseBlasTranspose::kNone        ? ASC_HIP_OP_N
         : t == SparseBlasTranspose::kTranspose ? ASC_HIP_OP_T
                                                : ASC_HIP_OP_H;
}
inline Result<CompletionEvent> Enqueued(const SparseHipContext& ctx, asc_hip_status s) {
  const Status st = ToStatus(s);
  if (!st.ok()) return st;
  return ctx.Record();
}

}  // namespace internal_sparse_hip

// Owner of a device CSR or CSC matrix (CudaCsrArray, providers/cuda.h:810-998):
// three buffers; view() is trusted.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipCompressedArray {
 public:
  HipCompressedArray() = default;
  [[nodiscard]] bool valid() const noexcept { return offsets_.data() != nullptr; }
  [[nodiscard]] extent_t rows() const noexcept { return rows_; }
  [[nodiscard]] extent_t columns() const noexcept { return cols_; }
  [[nodiscard]] nnz_t nnz() const noexcept { return nnz_; }
  [[nodiscard]] int index_bytes() const noexcept { return index_bytes_; }
  [[nodiscard]] Result<CompressedSparseView<Element, Format>> view() {
    if (!valid()) return Status(ErrorCode::kInvalidState, "HIP sparse array is empty");
    return HipProviderAccess::TrustedDevice<Element, Format>(
        offsets_.data(), indices_.data(), static_cast<Element*>(values_.data()), rows_, cols_, nnz_,
        index_bytes_);
  }
  [[nodiscard]] Result<CompressedSparseView<const Element, Format>> view() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "HIP sparse array is empty");
    return HipProviderAccess::TrustedDevice<const Element, Format>(
        offsets_.data(), indices_.data(), static_cast<const Element*>(values_.data()), rows_,
        cols_, nnz_, index_bytes_);
  }

  // Allocates the three buffers for a matrix of this shape.
  static Result<HipCompressedArray> Allocate(const HipMemoryResource& resource, extent_t rows,
                                             extent_t cols, nnz_t nnz, int index_bytes) {
    const extent_t outer = Format == SparseCompressedFormat::kCsr ? rows : cols;
    HipCompressedArray a;
    auto o = HipBuffer::Allocate(resource, sizeof(nnz_t) * static_cast<std::size_t>(outer + 1));
    if (!o.ok()) return o.status();
    auto i = HipBuffer::Allocate(resource, static_cast<std::size_t>(index_bytes) * nnz + 1);
    if (!i.ok()) return i.status();
    auto v = HipBuffer::Allocate(resource, sizeof(Element) * static_cast<std::size_t>(nnz) + 1);
    if (!v.ok()) return v.status();
    a.offsets_ = *o;
    a.indices_ = *i;
    a.values_ = *v;
    a.rows_ = rows;
    a.cols_ = cols;
    a.nnz_ = nnz;
    a.index_bytes_ = index_bytes;
    return a;
  }
  [[nodiscard]] void* offsets_data() const noexcept { return offsets_.data(); }
  [[nodiscard]] void* indices_data() const noexcept { return indices_.data(); }
  [[nodiscard]] void* values_data() const noexcept { return values_.data(); }

 private:
  HipBuffer offsets_, indices_, values_;
  extent_t rows_ = 0, cols_ = 0;
  nnz_t nnz_ = 0;
  int index_bytes_ = 8;
};

template <typename Element>
using HipCsrArray = HipCompressedArray<Element, SparseCompressedFormat::kCsr>;
template <typename Element>
using HipCscArray = HipCompressedArray<Element, SparseCompressedFormat::kCsc>;

// Array plus the completion of the work that fills it (CudaCsrClone,
// providers/cuda.h:1000-1038).
template <typename Element>
struct HipCsrClone {
  HipCsrArray<Element> array;
  CompletionEvent completion;
};
template <typename Element>
struct HipCscClone {
  HipCscArray<Element> array;
  CompletionEvent completion;
};

// Device coordinate (COO) matrix: interleaved int64 (row, col) and values.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipCoordinateArray {
 public:
  [[nodiscard]] bool valid() const noexcept { return coords_.data() != nullptr; }
  [[nodiscard]] Result<CoordinateView<const Element, 2>> view() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "HIP coordinate array is empty");
    return HipProviderAccess::TrustedDeviceCoordinates<const Element>(
        coords_.data(), static_cast<const Element*>(values_.data()), rows_, cols_, nnz_);
  }
  static Result<HipCoordinateArray> Allocate(const HipMemoryResource& resource, extent_t rows,
                                             extent_t cols, nnz_t nnz) {
    HipCoordinateArray a;
    auto c = HipBuffer::Allocate(resource, 16 * static_cast<std::size_t>(nnz) + 1);
    if (!c.ok()) return c.status();
    auto v = HipBuffer::Allocate(resource, sizeof(Element) * static_cast<std::size_t>(nnz) + 1);
    if (!v.ok()) return v.status();
    a.coords_ = *c;
    a.values_ = *v;
    a.rows_ = rows;
    a.cols_ = cols;
    a.nnz_ = nnz;
    return a;
  }
  [[nodiscard]] void* coords_data() const noexcept { return coords_.data(); }
  [[nodiscard]] void* values_data() const noexcept { return values_.data(); }
  // The entry count after a finalize that merged or dropped entries.
  void set_count(nnz_t nnz) noexcept { nnz_ = nnz; }

 private:
  HipBuffer coords_, values_;
  extent_t rows_ = 0, cols_ = 0;
  nnz_t nnz_ = 0;
};
template <typename Element>
struct HipCoordinateClone {
  HipCoordinateArray<Element> array;
  CompletionEvent completion;
};

// Index width of a clone: the reference's int64, or the explicit, validated
// narrowing to int32 (fails with kOverflow if an extent exceeds INT32_MAX).
enum class HipIndexWidth : std::uint8_t { k64 = 8, k32 = 4 };

// SpMV algorithm choice (ASC_HIP_ALG_*): kAuto uses the context's plan for the
// matrix (autotuned, learned or column-panel) and merge-path otherwise;
// kOrdered is the serial reference's summation, bit for bit.
enum class HipSpmvAlgorithm : std::uint8_t {
  kAuto = ASC_HIP_ALG_AUTO,
  kOrdered = ASC_HIP_ALG_ORDERED,
  kMergePath = ASC_HIP_ALG_MERGE,
  kVector1 = ASC_HIP_ALG_VECTOR1,
  kVector2 = ASC_HIP_ALG_VECTOR2,
  kVector4 = ASC_HIP_ALG_VECTOR4,
  kVector8 = ASC_HIP_ALG_VECTOR8,
  kVector16 = ASC_HIP_ALG_VECTOR16,
  kVector32 = ASC_HIP_ALG_VECTOR32,
  kVector64 = ASC_HIP_ALG_VECTOR64,
  kColumnPanel = ASC_HIP_ALG_PANEL,
};

// Host CSR -> device (CudaCloneCsr, providers/cuda.h:1040-1064).
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipCsrClone<std::remove_const_t<SourceElement>>> HipCloneCsr(
    SparseHipContext& context, CsrView<SourceElement> source, const HipMemoryResource& resource,
    HipIndexWidth width = HipIndexWidth::k64) {
  using E = std::remove_const_t<SourceElement>;
  if (resource.memory_space() != MemorySpace::kDevice)
    return Status(ErrorCode::kInvalidArgument, "a HIP clone needs a device memory resource");
  const asc_hip_sparse_desc src = HipProviderAccess::Describe(source);
  auto array = HipCsrArray<E>::Allocate(resource, source.rows(), source.columns(), source.nnz(),
                                        static_cast<int>(width));
  if (!array.ok()) return array.status();
  asc_hip_sparse_desc out{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_clone_compressed(
      context.native(), &src, static_cast<std::uint8_t>(width), array->offsets_data(),
      array->indices_data(), array->values_data(), &out));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCsrClone<E>{std::move(*array), *done};
}

// Host canonical COO -> device.
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipCoordinateClone<std::remove_const_t<SourceElement>>> HipCloneCoordinate(
    SparseHipContext& context, CoordinateView<SourceElement, 2> source,
    const HipMemoryResource& resource) {
  using E = std::remove_const_t<SourceElement>;
  auto array = HipCoordinateArray<E>::Allocate(resource, source.extent(0), source.extent(1),
                                               source.nnz());
  if (!array.ok()) return array.status();
  const asc_hip_sparse_desc src = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc out{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_clone_coordinate(
      context.native(), &src, array->coords_data(), array->values_data(), &out));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCoordinateClone<E>{std::move(*array), *done};
}

// ---- compat SpMV y = alpha*A*x + beta*y (blas.h:2031-2081) -----------------

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipCsrSpmvWorkspaceSize(
    SparseHipContext& context, CsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, HipStridedVectorView<Element> output,
    HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  const asc_hip_spmv_options opt{static_cast<std::uint8_t>(algorithm), ASC_HIP_ACC_PLAIN, {}};
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_csr_spmv_workspace_size(context.native(), &a, &x, &y, &opt, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipCsrSpmv(
    SparseHipContext& context, Element alpha, CsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, Element beta, HipStridedVectorView<Element> output,
    MutableMemoryView workspace, HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  const asc_hip_spmv_options opt{static_cast<std::uint8_t>(algorithm), ASC_HIP_ACC_PLAIN, {}};
  return internal_sparse_hip::Enqueued(
      context, asc_sparse_hip_csr_spmv(context.native(), static_cast<double>(alpha), &a, &x,
                                       static_cast<double>(beta), &y, workspace.data,
                                       workspace.size, &opt));
}

// Column-panel analysis of a device CSR matrix (ASC_HIP_ALG_PANEL); the
// analysis memory must outlive the plan.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipCsrSpmvAnalysisSize(SparseHipContext& context,
                                                         CsrView<const Element> matrix,
                                                         const asc_hip_panel_options* options = nullptr) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_csr_spmv_analysis_size(context.native(), &a, options, &bytes));
  if (!s.ok()) return s;
  return bytes;
}
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Status HipCsrSpmvAnalyze(SparseHipContext& context, CsrView<const Element> matrix,
                                       MutableMemoryView analysis,
                                       const asc_hip_panel_options* options = nullptr) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  return internal_sparse_hip::ToStatus(asc_sparse_hip_csr_spmv_analyze(
      context.native(), &a, options, analysis.data, analysis.size));
}

// ---- standardized SpMV y += alpha*op(A)*x (blas.h:1607-1667) ----------------

template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipSpmvWorkspaceSize(
    SparseHipContext& context, SparseBlasTranspose transpose,
    CompressedSparseView<const Element, Format> matrix, SparseBlasVectorView<const Element> input,
    SparseBlasVectorView<Element> output) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_standard_spmv_workspace_size(
      context.native(), internal_sparse_hip::Op(transpose), &a, &x, &y, nullptr, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

// CudaSpmv (providers/cuda.h:1616-1646).  CSR with kNone and CSC with a
// transpose need no workspace; the other two transpose into `workspace`.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSpmv(
    SparseHipContext& context, SparseBlasTranspose transpose, Element alpha,
    CompressedSparseView<const Element, Format> matrix, SparseBlasVectorView<const Element> input,
    SparseBlasVectorView<Element> output, MutableMemoryView workspace = {}) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  const asc_hip_scalar s{static_cast<double>(alpha), 0.0};
  return internal_sparse_hip::Enqueued(
      context, asc_sparse_hip_standard_spmv(context.native(), internal_sparse_hip::Op(transpose), s,
                                            &a, &x, &y, workspace.data, workspace.size, nullptr));
}

// ---- standardized SpMM C += alpha*op(A)*B (blas.h:1692-1750) ----------------

template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipSpmmWorkspaceSize(
    SparseHipContext& context, SparseBlasTranspose transpose,
    CompressedSparseView<const Element, Format> matrix, SparseBlasMatrixView<const Element> input,
    SparseBlasMatrixView<Element> output) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_matrix_desc b = internal_sparse_hip::Describe(input);
  const asc_hip_matrix_desc c = internal_sparse_hip::Describe(output);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_standard_spmm_workspace_size(
      context.native(), internal_sparse_hip::Op(transpose), &a, &b, &c, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

// CudaSpmm (providers/cuda.h:1648-1658).
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSpmm(
    SparseHipContext& context, SparseBlasTranspose transpose, Element alpha,
    CompressedSparseView<const Element, Format> matrix, SparseBlasMatrixView<const Element> input,
    SparseBlasMatrixView<Element> output, MutableMemoryView workspace = {}) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(matrix);
  const asc_hip_matrix_desc b = internal_sparse_hip::Describe(input);
  const asc_hip_matrix_desc c = internal_sparse_hip::Describe(output);
  const asc_hip_scalar s{static_cast<double>(alpha), 0.0};
  return internal_sparse_hip::Enqueued(
      context, asc_sparse_hip_standard_spmm(context.native(), internal_sparse_hip::Op(transpose), s,
                                            &a, &b, &c, workspace.data, workspace.size));
}

// ---- conversions (compressed.h:1675-1837) ----------------------------------

// CSR -> CSC (ConvertToCsc(ctx, CsrView), compressed.h:1753-1760) and CSC ->
// CSR (:1727-1734): a stable device transpose, bit-exact with the reference.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipTransposeWorkspaceSize(
    SparseHipContext& context, CompressedSparseView<const Element, Format> source) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_convert_compressed_workspace_size(context.native(), &a, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCscClone<Element>> HipConvertToCsc(SparseHipContext& context,
                                                           CsrView<const Element> source,
                                                           const HipMemoryResource& resource,
                                                           MutableMemoryView workspace) {
  auto out = HipCscArray<Element>::Allocate(resource, source.rows(), source.columns(),
                                            source.nnz(), source.index_bytes());
  if (!out.ok()) return out.status();
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_compressed(
      context.native(), &a, out->offsets_data(), out->indices_data(), out->values_data(), &d,
      workspace.data, workspace.size));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCscClone<Element>{std::move(*out), *done};
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCsrClone<Element>> HipConvertToCsr(SparseHipContext& context,
                                                           CscView<const Element> source,
                                                           const HipMemoryResource& resource,
                                                           MutableMemoryView workspace) {
  auto out = HipCsrArray<Element>::Allocate(resource, source.rows(), source.columns(),
                                            source.nnz(), source.index_bytes());
  if (!out.ok()) return out.status();
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_compressed(
      context.native(), &a, out->offsets_data(), out->indices_data(), out->values_data(), &d,
      workspace.data, workspace.size));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCsrClone<Element>{std::move(*out), *done};
}

// Canonical device COO -> CSR / CSC (ConvertToCsr/ConvertToCsc(ctx,
// CoordinateView), compressed.h:1675-1708).
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipConvertCoordinateWorkspaceSize(
    SparseHipContext& context, CoordinateView<const Element, 2> source,
    SparseCompressedFormat format, HipIndexWidth width = HipIndexWidth::k64) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_coordinate_workspace_size(
      context.native(), &a, format == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC,
      static_cast<std::uint8_t>(width), &bytes));
  if (!s.ok()) return s;
  return bytes;
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCsrClone<Element>> HipConvertToCsr(
    SparseHipContext& context, CoordinateView<const Element, 2> source,
    const HipMemoryResource& resource, MutableMemoryView workspace,
    HipIndexWidth width = HipIndexWidth::k64) {
  auto out = HipCsrArray<Element>::Allocate(resource, source.extent(0), source.extent(1),
                                            source.nnz(), static_cast<int>(width));
  if (!out.ok()) return out.status();
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_coordinate(
      context.native(), &a, ASC_HIP_CSR, static_cast<std::uint8_t>(width), out->offsets_data(),
      out->indices_data(), out->values_data(), &d, workspace.data, workspace.size));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCsrClone<Element>{std::move(*out), *done};
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCscClone<Element>> HipConvertToCsc(
    SparseHipContext& context, CoordinateView<const Element, 2> source,
    const HipMemoryResource& resource, MutableMemoryView workspace,
    HipIndexWidth width = HipIndexWidth::k64) {
  auto out = HipCscArray<Element>::Allocate(resource, source.extent(0), source.extent(1),
                                            source.nnz(), static_cast<int>(width));
  if (!out.ok()) return out.status();
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_coordinate(
      context.native(), &a, ASC_HIP_CSC, static_cast<std::uint8_t>(width), out->offsets_data(),
      out->indices_data(), out->values_data(), &d, workspace.data, workspace.size));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCscClone<Element>{std::move(*out), *done};
}

// Device CSR/CSC -> canonical COO (ConvertToCoordinate, compressed.h:1781-1837).
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipConvertToCoordinateWorkspaceSize(
    SparseHipContext& context, CompressedSparseView<const Element, Format> source) {
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_convert_to_coordinate_workspace_size(context.native(), &a, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCoordinateClone<Element>> HipConvertToCoordinate(
    SparseHipContext& context, CompressedSparseView<const Element, Format> source,
    const HipMemoryResource& resource, MutableMemoryView workspace = {}) {
  auto out = HipCoordinateArray<Element>::Allocate(resource, source.rows(), source.columns(),
                                                   source.nnz());
  if (!out.ok()) return out.status();
  const asc_hip_sparse_desc a = HipProviderAccess::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_to_coordinate(
      context.native(), &a, out->coords_data(), out->values_data(), &d, workspace.data,
      workspace.size));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipCoordinateClone<Element>{std::move(*out), *done};
}

// Device CoordinateBuilder::Finalize (coordinate.h:1300-1497): raw device
// triplets (interleaved int64 (row, col) in insertion order, values) to a
// canonical coordinate array.  DuplicatePolicy / ExplicitZeroPolicy values
// follow coordinate.h:60-78 (ASC_HIP_DUPLICATE_*, ASC_HIP_ZERO_*).  Waits for
// the device; the array's view() has the finalized count.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCoordinateArray<Element>> HipFinalizeCoordinates(
    SparseHipContext& context, const index_t* device_coordinates, const Element* device_values,
    extent_t rows, extent_t cols, nnz_t count, std::uint8_t duplicate_policy,
    std::uint8_t zero_policy, const HipMemoryResource& resource) {
  asc_hip_sparse_desc t{};
  t.outer_offsets = device_coordinates;
  t.values = device_values;
  t.rows = rows;
  t.cols = cols;
  t.nnz = count;
  t.format = ASC_HIP_COO;
  t.element_kind = internal_sparse_hip::ElementKind<Element>();
  t.index_bytes = 8;
  t.offset_bytes = 8;
  t.memory_space = ASC_HIP_DEVICE;
  std::size_t bytes = 0;
  Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_coo_finalize_workspace_size(context.native(), &t, &bytes));
  if (!s.ok()) return s;
  auto ws = HipBuffer::Allocate(resource, bytes);
  if (!ws.ok()) return ws.status();
  auto out = HipCoordinateArray<Element>::Allocate(resource, rows, cols, count);
  if (!out.ok()) return out.status();
  asc_hip_sparse_desc d{};
  s = internal_sparse_hip::ToStatus(asc_sparse_hip_coo_finalize(
      context.native(), &t, duplicate_policy, zero_policy, out->coords_data(), out->values_data(),
      &d, ws->data(), ws->size()));
  if (!s.ok()) return s;
  out->set_count(d.nnz);
  return std::move(*out);
}

// ---- sparse Level-1 (blas.h:1330-1582; CudaSparseDot ... CudaSparseScatter,
// providers/cuda.h:1464-1614) ----------------------------------------------

// Owner of a device indexed vector (CudaIndexedVectorArray,
// providers/cuda.h:1066-1220); view() is trusted.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipIndexedVectorArray {
 public:
  [[nodiscard]] bool valid() const noexcept { return indices_.data() != nullptr; }
  [[nodiscard]] Result<SparseBlasIndexedVectorView<Element>> view() {
    if (!valid()) return Status(ErrorCode::kInvalidState, "HIP indexed vector is empty");
    return HipProviderAccess::TrustedDeviceIndexed<Element>(
        indices_.data(), static_cast<Element*>(values_.data()), nnz_, extent_);
  }
  [[nodiscard]] Result<SparseBlasIndexedVectorView<const Element>> view() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "HIP indexed vector is empty");
    return HipProviderAccess::TrustedDeviceIndexed<const Element>(
        indices_.data(), static_cast<const Element*>(values_.data()), nnz_, extent_);
  }
  static Result<HipIndexedVectorArray> Allocate(const HipMemoryResource& resource, nnz_t nnz,
                                                extent_t extent) {
    HipIndexedVectorArray a;
    auto i = HipBuffer::Allocate(resource, 8 * static_cast<std::size_t>(nnz) + 8);
    if (!i.ok()) return i.status();
    auto v = HipBuffer::Allocate(resource, sizeof(Element) * static_cast<std::size_t>(nnz) + 8);
    if (!v.ok()) return v.status();
    a.indices_ = *i;
    a.values_ = *v;
    a.nnz_ = nnz;
    a.extent_ = extent;
    return a;
  }
  [[nodiscard]] void* indices_data() const noexcept { return indices_.data(); }
  [[nodiscard]] void* values_data() const noexcept { return values_.data(); }

 private:
  HipBuffer indices_, values_;
  nnz_t nnz_ = 0;
  extent_t extent_ = 0;
};

template <typename Element>
struct HipIndexedVectorClone {
  HipIndexedVectorArray<Element> array;
  CompletionEvent completion;
};

// Host canonical indexed vector -> device (CudaCloneIndexedVector,
// providers/cuda.h:1247-1272).
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipIndexedVectorClone<std::remove_const_t<SourceElement>>>
HipCloneIndexedVector(SparseHipContext& context, SparseBlasIndexedVectorView<SourceElement> source,
                      const HipMemoryResource& resource) {
  using E = std::remove_const_t<SourceElement>;
  auto array = HipIndexedVectorArray<E>::Allocate(resource, source.nnz(), source.dense_extent());
  if (!array.ok()) return array.status();
  const asc_hip_indexed_desc src = HipProviderAccess::Describe(source);
  asc_hip_indexed_desc out{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_clone_indexed(
      context.native(), &src, array->indices_data(), array->values_data(), &out));
  if (!s.ok()) return s;
  auto done = context.Record();
  if (!done.ok()) return done.status();
  return HipIndexedVectorClone<E>{std::move(*array), *done};
}

namespace internal_sparse_hip {

template <typename Element>
Result<CompletionEvent> Level1(SparseHipContext& context, std::uint8_t op,
                               SparseBlasConjugation conj, Element alpha,
                               const asc_hip_indexed_desc& sparse, const asc_hip_vector_desc& dense,
                               const asc_hip_vector_desc* result, HipSpmvAlgorithm algorithm) {
  const asc_hip_spmv_options opt{static_cast<std::uint8_t>(algorithm), ASC_HIP_ACC_PLAIN, {}};
  const asc_hip_scalar a{static_cast<double>(alpha), 0.0};
  return Enqueued(context, asc_sparse_hip_standard_level1(
                               context.native(), op,
                               conj == SparseBlasConjugation::kConjugated ? ASC_HIP_CONJUGATED
                                                                          : ASC_HIP_UNCONJUGATED,
                               a, &sparse, &dense, result, &opt));
}

}  // namespace internal_sparse_hip

// Writes the dot product into `result` (one element).  kOrdered: the serial
// reference's left-to-right sum, bit for bit; kAuto: a deterministic tree.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseDot(
    SparseHipContext& context, SparseBlasConjugation conjugation,
    SparseBlasIndexedVectorView<const Element> sparse, SparseBlasVectorView<const Element> dense,
    SparseBlasVectorView<Element> result, HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  const auto r = internal_sparse_hip::Describe(result);
  return internal_sparse_hip::Level1<Element>(context, ASC_HIP_LEVEL1_DOT, conjugation, Element(1),
                                              HipProviderAccess::Describe(sparse),
                                              internal_sparse_hip::Describe(dense), &r, algorithm);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseAxpy(SparseHipContext& context, Element alpha,
                                                    SparseBlasIndexedVectorView<const Element> sparse,
                                                    SparseBlasVectorView<Element> dense) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_AXPY, SparseBlasConjugation::kUnconjugated, alpha,
      HipProviderAccess::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseGather(SparseHipContext& context,
                                                      SparseBlasVectorView<const Element> dense,
                                                      SparseBlasIndexedVectorView<Element> sparse) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_GATHER, SparseBlasConjugation::kUnconjugated, Element(1),
      HipProviderAccess::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseGatherZero(SparseHipContext& context,
                                                          SparseBlasVectorView<Element> dense,
                                                          SparseBlasIndexedVectorView<Element> sparse) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_GATHER_ZERO, SparseBlasConjugation::kUnconjugated, Element(1),
      HipProviderAccess::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseScatter(SparseHipContext& context,
                                                       SparseBlasIndexedVectorView<const Element> sparse,
                                                       SparseBlasVectorView<Element> dense) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_SCATTER, SparseBlasConjugation::kUnconjugated, Element(1),
      HipProviderAccess::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

// ---- triangular solve (blas.h:1760-1880; CudaSparseTriangularSolve,
// providers/cuda.h:1680-1756) -------------------------------------------------

// A triangular matrix with its level schedule for one operation: the device
// form of SparseBlasTriangularView (blas.h:1074-1160) plus the analysis.
// Analyze checks what the view's Create checks; the matrix's arrays and the
// analysis must stay unmodified while the plan is used.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipTriangularPlan {
 public:
  static Result<HipTriangularPlan> Analyze(SparseHipContext& context,
                                           CompressedSparseView<const Element, Format> matrix,
                                           SparseBlasTriangle triangle, SparseBlasDiagonal diagonal,
                                           SparseBlasTranspose transpose,
                                           const HipMemoryResource& resource) {
    HipTriangularPlan p;
    p.desc_ = HipProviderAccess::Describe(matrix);
    p.transpose_ = internal_sparse_hip::Op(transpose);
    std::size_t bytes = 0;
    Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_triangular_analysis_size(
        context.native(), p.transpose_, &p.desc_, &bytes));
    if (!s.ok()) return s;
    auto buf = HipBuffer::Allocate(resource, bytes);
    if (!buf.ok()) return buf.status();
    p.analysis_ = *buf;
    s = internal_sparse_hip::ToStatus(asc_sparse_hip_triangular_analyze(
        context.native(), p.transpose_,
        triangle == SparseBlasTriangle::kLower ? ASC_HIP_LOWER : ASC_HIP_UPPER,
        diagonal == SparseBlasDiagonal::kUnit ? ASC_HIP_UNIT : ASC_HIP_NON_UNIT, &p.desc_,
        p.analysis_.data(), p.analysis_.size(), &p.info_));
    if (!s.ok()) return s;
    return p;
  }
  [[nodiscard]] std::int64_t levels() const noexcept { return info_.levels; }

  // right_hand_sides := alpha * op(A)^-1 * right_hand_sides
  // (SparseTriangularSolveMultiple; a vector is the one-column case).
  [[nodiscard]] Result<CompletionEvent> Solve(SparseHipContext& context, Element alpha,
                                              SparseBlasMatrixView<Element> right_hand_sides) const {
    const asc_hip_matrix_desc b = internal_sparse_hip::Describe(right_hand_sides);
    const asc_hip_scalar a{static_cast<double>(alpha), 0.0};
    return internal_sparse_hip::Enqueued(
        context, asc_sparse_hip_triangular_solve(context.native(), transpose_, a, &desc_, &b,
                                                 analysis_.data(), analysis_.size(), &info_));
  }
  [[nodiscard]] Result<CompletionEvent> Solve(SparseHipContext& context, Element alpha,
                                              SparseBlasVectorView<Element> right_hand_side) const {
    if (right_hand_side.increment() <= 0)
      return Status(ErrorCode::kUnsupported, "the device solve takes a positive increment");
    auto m = SparseBlasMatrixView<Element>::Create(
        right_hand_side.data(), right_hand_side.size(), 1, SparseBlasLayout::kRowMajor,
        right_hand_side.increment(),
        {right_hand_side.data(),
         sizeof(Element) * static_cast<std::size_t>(right_hand_side.size() > 0
                                                        ? (right_hand_side.size() - 1) *
                                                                  right_hand_side.increment() +
                                                              1
                                                        : 0),
         right_hand_side.memory_space()});
    if (!m.ok()) return m.status();
    return Solve(context, alpha, *m);
  }

 private:
  asc_hip_sparse_desc desc_{};
  asc_hip_triangular_info info_{};
  std::uint8_t transpose_ = ASC_HIP_OP_N;
  HipBuffer analysis_;
};

// Stream-ordered copies between host and device memory (CopyBytes,
// src/core/cuda/runtime.cc:226-253).
[[nodiscard]] inline Result<CompletionEvent> HipCopy(SparseHipContext& context, void* dst,
                                                     const void* src, std::size_t bytes) {
  return internal_sparse_hip::Enqueued(context,
                                       asc_sparse_hip_copy(context.native(), dst, src, bytes));
}

}  // namespace asc
