// ASC::sparse_hip: the C++20 facet of the MI355X Sparse BLAS provider.
//
// The public surface mirrors the reference's CUDA facet
// (include/asc/sparse/providers/cuda.h) with Cuda -> Hip, on the reference's
// own core types: SparseHipContext::Create(ExecutionContext) (:499-606) over a
// HIP ExecutionContext from CreateHipExecutionContext (core_hip.hpp),
// HipStridedVectorView (:616-730), HipCsrArray / HipCsrClone /
// HipCloneCsr(context, CsrView, MemoryResource&) (:810-1056),
// HipCsrSpmvWorkspaceSize + HipCsrSpmv (:1758-1802), HipSpmv / HipSpmm
// (:1616-1658), HipCloneIndexedVector and the Level-1 operations
// (:1066-1614), each returning Result<...> and never throwing.  It adds the
// device conversions and finalize the reference provider lacks, the
// column-panel analysis, a level-scheduled triangular plan, and an explicit
// 32-bit column-index clone (HipIndexWidth::k32, viewed through the
// facet-side HipNarrowCompressedView, since the reference's views hold
// 64-bit indices).
//
// Every operation packs plain descriptors and calls the C ABI of
// include/asc_sparse_hip.h, the type-erased seam that replaces the
// reference's *Erased functions (providers/cuda.h:212-253).  Failures come
// back as Status with the asc::ErrorCode numbering, provider "hip" and the
// hipError_t as native code.  Enqueue success does not imply completion: wait
// on the returned CompletionEvent.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <type_traits>
#include <utility>

#include "asc_sparse_hip.h"
#include "asc_sparse_hip/core_hip.hpp"
#include "asc_sparse_hip/vocabulary.hpp"

namespace asc {

class SparseHipContext;

namespace internal_sparse_hip {

inline Status ToStatus(asc_hip_status s) {
  if (s.code == ASC_HIP_OK) return Status();
  return Status(static_cast<ErrorCode>(s.code), asc_sparse_hip_last_message(),
                s.provider == ASC_HIP_PROVIDER_HIP ? "hip" : "", s.native);
}

// Element kinds the device kernels compute in (the CUDA facet's
// kSupportedElement, providers/cuda.h:190-193).
template <typename E>
inline constexpr bool kSupportedElement =
    !std::is_volatile_v<E> && (std::is_same_v<std::remove_cv_t<E>, float> ||
                               std::is_same_v<std::remove_cv_t<E>, double>);

template <typename E>
constexpr std::uint8_t ElementKind() {
  return std::is_same_v<std::remove_cv_t<E>, double> ? ASC_HIP_F64 : ASC_HIP_F32;
}

struct ContextDeleter {
  void operator()(asc_hip_context* c) const noexcept { (void)asc_sparse_hip_context_destroy(c); }
};

inline MutableMemoryView NoWorkspace() noexcept {
  return MutableMemoryView(nullptr, 0, MemorySpace::kDevice);
}

class Access;

}  // namespace internal_sparse_hip

// Provider state bound to a HIP execution context (SparseCudaContext,
// providers/cuda.h:499-606; context.cc:69-113): the library context borrows
// the execution context's stream, so work of both is ordered.  Move-only;
// one context must not be used from two threads at once.
class SparseHipContext {
 public:
  static Result<SparseHipContext> Create(ExecutionContext execution_context);
  // A context that validates every call and enqueues nothing (no device is
  // touched): calls that pass validation return kUnavailable.  Facet
  // extension for testing rejection paths without a GPU.
  static Result<SparseHipContext> CreateValidationOnly();
  SparseHipContext(const SparseHipContext&) = delete;
  SparseHipContext& operator=(const SparseHipContext&) = delete;
  SparseHipContext(SparseHipContext&& other) noexcept = default;
  SparseHipContext& operator=(SparseHipContext&& other) noexcept = default;
  ~SparseHipContext() = default;

  [[nodiscard]] const ExecutionContext& execution_context() const noexcept {
    return execution_context_;
  }
  // The C ABI context (facet extension).
  [[nodiscard]] asc_hip_context* native() const noexcept { return state_.get(); }
  // Waits for the stream; reports a kernel that gave up a bounded wait.
  [[nodiscard]] Status Synchronize() const {
    return internal_sparse_hip::ToStatus(asc_sparse_hip_context_synchronize(state_.get()));
  }

 private:
  SparseHipContext(ExecutionContext execution_context, asc_hip_context* state) noexcept
      : execution_context_(std::move(execution_context)), state_(state) {}
  ExecutionContext execution_context_;
  std::unique_ptr<asc_hip_context, internal_sparse_hip::ContextDeleter> state_;
};

namespace internal_sparse_hip {

// Completion of the work a successful call enqueued.
inline Result<CompletionEvent> Enqueued(SparseHipContext& context, asc_hip_status s) {
  const Status st = ToStatus(s);
  if (!st.ok()) return st;
  return RecordHipEvent(context.execution_context());
}

inline Result<Buffer> DeviceBuffer(MemoryResource& resource, std::size_t bytes) {
  if (resource.space() != MemorySpace::kDevice)
    return Status(ErrorCode::kInvalidArgument, "HIP device arrays need a device memory resource");
  return Buffer::Allocate(resource, bytes, 256);
}

}  // namespace internal_sparse_hip

// HIP-accessible strided vector (CudaStridedVectorView, providers/cuda.h:616-730).
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipStridedVectorView {
 public:
  static Result<HipStridedVectorView> Create(Element* data, extent_t extent, stride_t stride) {
    if (extent < 0 || stride <= 0)
      return Status(ErrorCode::kInvalidArgument,
                    "A HIP strided vector requires nonnegative extent and positive stride");
    if (extent != 0 && data == nullptr)
      return Status(ErrorCode::kInvalidArgument, "A nonempty HIP strided vector cannot be null");
    return HipStridedVectorView(data, extent, stride);
  }
  template <typename OtherElement>
    requires(std::is_const_v<Element> &&
             std::same_as<std::remove_const_t<OtherElement>, std::remove_const_t<Element>> &&
             !std::is_const_v<OtherElement>)
  constexpr HipStridedVectorView(HipStridedVectorView<OtherElement> other) noexcept  // NOLINT
      : data_(other.data()), extent_(other.extent()), stride_(other.stride()) {}
  [[nodiscard]] constexpr Element* data() const noexcept { return data_; }
  [[nodiscard]] constexpr extent_t extent() const noexcept { return extent_; }
  [[nodiscard]] constexpr stride_t stride() const noexcept { return stride_; }
  [[nodiscard]] static constexpr MemorySpace memory_space() noexcept { return MemorySpace::kDevice; }

 private:
  constexpr HipStridedVectorView(Element* data, extent_t extent, stride_t stride) noexcept
      : data_(data), extent_(extent), stride_(stride) {}
  Element* data_;
  extent_t extent_;
  stride_t stride_;
};

// Column-index width of a device clone: the reference's int64, or the
// explicit, validated narrowing to int32 (kOverflow when the inner extent
// exceeds INT32_MAX).  Facet extension: narrow clones halve the index bytes
// every SpMV streams.
enum class HipIndexWidth : std::uint8_t { k64 = 8, k32 = 4 };

// A trusted device CSR/CSC view over 32-bit inner indices, made only by this
// facet from a HipIndexWidth::k32 clone (facet extension; the reference's
// CompressedSparseView holds 64-bit indices).
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipNarrowCompressedView {
 public:
  using value_type = std::remove_const_t<Element>;
  template <typename OtherElement>
    requires(std::is_const_v<Element> && std::same_as<std::remove_const_t<OtherElement>, value_type> &&
             !std::is_const_v<OtherElement>)
  constexpr HipNarrowCompressedView(const HipNarrowCompressedView<OtherElement, Format>& o) noexcept  // NOLINT
      : offsets_(o.outer_offsets()), indices_(o.inner_indices()), values_(o.values()),
        shape_(o.shape()), nonzeros_(o.nnz()) {}
  [[nodiscard]] constexpr const std::array<extent_t, 2>& shape() const noexcept { return shape_; }
  [[nodiscard]] constexpr extent_t rows() const noexcept { return shape_[0]; }
  [[nodiscard]] constexpr extent_t columns() const noexcept { return shape_[1]; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr const nnz_t* outer_offsets() const noexcept { return offsets_; }
  [[nodiscard]] constexpr const std::int32_t* inner_indices() const noexcept { return indices_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] static constexpr MemorySpace memory_space() noexcept { return MemorySpace::kDevice; }
  [[nodiscard]] static constexpr bool canonical_structure_trusted() noexcept { return true; }

 private:
  template <typename E2, SparseCompressedFormat F2>
    requires internal_sparse_hip::kSupportedElement<E2>
  friend class HipNarrowCompressedView;
  friend class internal_sparse_hip::Access;
  constexpr HipNarrowCompressedView(const nnz_t* offsets, const std::int32_t* indices,
                                    Element* values, std::array<extent_t, 2> shape,
                                    nnz_t nonzeros) noexcept
      : offsets_(offsets), indices_(indices), values_(values), shape_(shape), nonzeros_(nonzeros) {}
  const nnz_t* offsets_;
  const std::int32_t* indices_;
  Element* values_;
  std::array<extent_t, 2> shape_;
  nnz_t nonzeros_;
};

template <typename Element>
using HipNarrowCsrView = HipNarrowCompressedView<Element, SparseCompressedFormat::kCsr>;
template <typename Element>
using HipNarrowCscView = HipNarrowCompressedView<Element, SparseCompressedFormat::kCsc>;

// Owner of a device CSR or CSC matrix (CudaCsrArray, providers/cuda.h:810-998):
// three buffers from the caller's MemoryResource; view() is trusted.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipCompressedArray {
 public:
  HipCompressedArray(const HipCompressedArray&) = delete;
  HipCompressedArray& operator=(const HipCompressedArray&) = delete;
  HipCompressedArray(HipCompressedArray&&) noexcept = default;
  HipCompressedArray& operator=(HipCompressedArray&&) noexcept = default;
  ~HipCompressedArray() = default;

  [[nodiscard]] bool valid() const noexcept {
    return outer_offsets_.valid() && inner_indices_.valid() && values_.valid();
  }
  [[nodiscard]] extent_t rows() const noexcept { return shape_[0]; }
  [[nodiscard]] extent_t columns() const noexcept { return shape_[1]; }
  [[nodiscard]] nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] HipIndexWidth index_width() const noexcept { return width_; }

  [[nodiscard]] Result<CompressedSparseView<Element, Format>> view() { return View<Element>(); }
  [[nodiscard]] Result<CompressedSparseView<const Element, Format>> view() const {
    return View<const Element>();
  }
  [[nodiscard]] Result<HipNarrowCompressedView<Element, Format>> narrow_view() {
    return Narrow<Element>();
  }
  [[nodiscard]] Result<HipNarrowCompressedView<const Element, Format>> narrow_view() const {
    return Narrow<const Element>();
  }

 private:
  friend class internal_sparse_hip::Access;
  HipCompressedArray(std::array<extent_t, 2> shape, nnz_t nonzeros, HipIndexWidth width,
                     Buffer offsets, Buffer indices, Buffer values) noexcept
      : shape_(shape), nonzeros_(nonzeros), width_(width), outer_offsets_(std::move(offsets)),
        inner_indices_(std::move(indices)), values_(std::move(values)) {}
  template <typename E>
  Result<CompressedSparseView<E, Format>> View() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "A moved-from HIP sparse array has no view");
    if (width_ != HipIndexWidth::k64)
      return Status(ErrorCode::kInvalidState,
                    "A 32-bit-index HIP clone is viewed through narrow_view()");
    return internal_sparse_compressed::ProviderAccess::MakeCanonicalView<E, Format>(
        static_cast<const nnz_t*>(outer_offsets_.data()),
        static_cast<const index_t*>(inner_indices_.data()),
        static_cast<E*>(const_cast<void*>(values_.data())), shape_, nonzeros_, MemorySpace::kDevice);
  }
  template <typename E>
  Result<HipNarrowCompressedView<E, Format>> Narrow() const;

  std::array<extent_t, 2> shape_;
  nnz_t nonzeros_;
  HipIndexWidth width_;
  Buffer outer_offsets_;
  Buffer inner_indices_;
  Buffer values_;
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

// Device canonical coordinate (COO) matrix: interleaved int64 (row, col)
// pairs and values; view() is trusted.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipCoordinateArray {
 public:
  HipCoordinateArray(const HipCoordinateArray&) = delete;
  HipCoordinateArray& operator=(const HipCoordinateArray&) = delete;
  HipCoordinateArray(HipCoordinateArray&&) noexcept = default;
  HipCoordinateArray& operator=(HipCoordinateArray&&) noexcept = default;
  ~HipCoordinateArray() = default;
  [[nodiscard]] bool valid() const noexcept { return coordinates_.valid() && values_.valid(); }
  [[nodiscard]] nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] Result<CoordinateView<const Element, 2>> view() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "A moved-from HIP coordinate array has no view");
    return internal_sparse_coordinate::ProviderAccess::MakeCanonicalView<const Element, 2>(
        static_cast<const index_t*>(coordinates_.data()),
        static_cast<const Element*>(values_.data()), extents_, nonzeros_, MemorySpace::kDevice);
  }

 private:
  friend class internal_sparse_hip::Access;
  HipCoordinateArray(std::array<extent_t, 2> extents, nnz_t nonzeros, Buffer coordinates,
                     Buffer values) noexcept
      : extents_(extents), nonzeros_(nonzeros), coordinates_(std::move(coordinates)),
        values_(std::move(values)) {}
  std::array<extent_t, 2> extents_;
  nnz_t nonzeros_;
  Buffer coordinates_;
  Buffer values_;
};
template <typename Element>
struct HipCoordinateClone {
  HipCoordinateArray<Element> array;
  CompletionEvent completion;
};

// Owner of a device indexed vector (CudaIndexedVectorArray,
// providers/cuda.h:1066-1220); view() is trusted.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
class HipIndexedVectorArray {
 public:
  HipIndexedVectorArray(const HipIndexedVectorArray&) = delete;
  HipIndexedVectorArray& operator=(const HipIndexedVectorArray&) = delete;
  HipIndexedVectorArray(HipIndexedVectorArray&&) noexcept = default;
  HipIndexedVectorArray& operator=(HipIndexedVectorArray&&) noexcept = default;
  ~HipIndexedVectorArray() = default;
  [[nodiscard]] bool valid() const noexcept { return indices_.valid() && values_.valid(); }
  [[nodiscard]] Result<SparseBlasIndexedVectorView<Element>> view() { return View<Element>(); }
  [[nodiscard]] Result<SparseBlasIndexedVectorView<const Element>> view() const {
    return View<const Element>();
  }

 private:
  friend class internal_sparse_hip::Access;
  HipIndexedVectorArray(extent_t dense_extent, nnz_t nonzeros, Buffer indices, Buffer values) noexcept
      : dense_extent_(dense_extent), nonzeros_(nonzeros), indices_(std::move(indices)),
        values_(std::move(values)) {}
  template <typename E>
  Result<SparseBlasIndexedVectorView<E>> View() const {
    if (!valid()) return Status(ErrorCode::kInvalidState, "A moved-from HIP indexed vector has no view");
    return internal_sparse_standard_blas::ProviderAccess::MakeCanonicalIndexedVector<E>(
        static_cast<const index_t*>(indices_.data()),
        static_cast<E*>(const_cast<void*>(values_.data())), nonzeros_, dense_extent_,
        ConstMemoryView(indices_.data(), indices_.size(), MemorySpace::kDevice),
        ConstMemoryView(values_.data(), values_.size(), MemorySpace::kDevice));
  }
  extent_t dense_extent_;
  nnz_t nonzeros_;
  Buffer indices_;
  Buffer values_;
};
template <typename Element>
struct HipIndexedVectorClone {
  HipIndexedVectorArray<Element> array;
  CompletionEvent completion;
};

namespace internal_sparse_hip {

// Builds facet arrays and views and describes operands for the C ABI.
class Access {
 public:
  template <typename E, SparseCompressedFormat F>
  static HipCompressedArray<E, F> MakeCompressed(std::array<extent_t, 2> shape, nnz_t nnz,
                                                 HipIndexWidth w, Buffer o, Buffer i, Buffer v) {
    return HipCompressedArray<E, F>(shape, nnz, w, std::move(o), std::move(i), std::move(v));
  }
  template <typename E>
  static HipCoordinateArray<E> MakeCoordinate(std::array<extent_t, 2> extents, nnz_t nnz, Buffer c,
                                              Buffer v) {
    return HipCoordinateArray<E>(extents, nnz, std::move(c), std::move(v));
  }
  template <typename E>
  static HipIndexedVectorArray<E> MakeIndexed(extent_t extent, nnz_t nnz, Buffer i, Buffer v) {
    return HipIndexedVectorArray<E>(extent, nnz, std::move(i), std::move(v));
  }
  template <typename E, SparseCompressedFormat F>
  static HipNarrowCompressedView<E, F> MakeNarrow(const nnz_t* o, const std::int32_t* i, E* v,
                                                  std::array<extent_t, 2> shape, nnz_t nnz) {
    return HipNarrowCompressedView<E, F>(o, i, v, shape, nnz);
  }
};

template <typename E, SparseCompressedFormat F>
asc_hip_sparse_desc Describe(const CompressedSparseView<E, F>& v) {
  asc_hip_sparse_desc d{};
  d.outer_offsets = v.outer_offsets();
  d.inner_indices = v.inner_indices();
  d.values = v.values();
  d.rows = v.rows();
  d.cols = v.columns();
  d.nnz = v.nnz();
  d.format = F == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC;
  d.element_kind = ElementKind<E>();
  d.index_bytes = 8;
  d.offset_bytes = 8;
  d.memory_space = static_cast<std::uint8_t>(v.memory_space());
  d.trusted = v.canonical_structure_trusted() ? 1 : 0;
  return d;
}
template <typename E, SparseCompressedFormat F>
asc_hip_sparse_desc Describe(const HipNarrowCompressedView<E, F>& v) {
  asc_hip_sparse_desc d{};
  d.outer_offsets = v.outer_offsets();
  d.inner_indices = v.inner_indices();
  d.values = v.values();
  d.rows = v.rows();
  d.cols = v.columns();
  d.nnz = v.nnz();
  d.format = F == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC;
  d.element_kind = ElementKind<E>();
  d.index_bytes = 4;
  d.offset_bytes = 8;
  d.memory_space = ASC_HIP_DEVICE;
  d.trusted = 1;
  return d;
}
template <typename E>
asc_hip_sparse_desc Describe(const CoordinateView<E, 2>& v) {
  asc_hip_sparse_desc d{};
  d.outer_offsets = v.coordinates();
  d.values = v.values();
  d.rows = v.shape()[0];
  d.cols = v.shape()[1];
  d.nnz = v.nnz();
  d.format = ASC_HIP_COO;
  d.element_kind = ElementKind<E>();
  d.index_bytes = 8;
  d.offset_bytes = 8;
  d.memory_space = static_cast<std::uint8_t>(v.memory_space());
  d.trusted = v.canonical_structure_trusted() ? 1 : 0;
  return d;
}
template <typename E>
asc_hip_indexed_desc Describe(const SparseBlasIndexedVectorView<E>& v) {
  asc_hip_indexed_desc d{};
  d.indices = v.indices();
  d.values = v.values();
  d.nnz = v.nnz();
  d.dense_extent = v.dense_extent();
  d.element_kind = ElementKind<E>();
  d.memory_space = static_cast<std::uint8_t>(v.memory_space());
  d.writable = std::is_const_v<E> ? 0 : 1;
  d.trusted = v.canonical_structure_trusted() ? 1 : 0;
  return d;
}
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

// Three device buffers for a compressed matrix of this shape, then the C ABI
// call that fills them.
template <typename E, SparseCompressedFormat F, typename Fill>
Result<HipCompressedArray<E, F>> FillCompressed(MemoryResource& resource, extent_t rows,
                                                extent_t cols, nnz_t nnz, HipIndexWidth width,
                                                Fill&& fill) {
  const extent_t outer = F == SparseCompressedFormat::kCsr ? rows : cols;
  auto o = DeviceBuffer(resource, sizeof(nnz_t) * static_cast<std::size_t>(outer + 1));
  if (!o.ok()) return o.status();
  auto i = DeviceBuffer(resource, static_cast<std::size_t>(width) * static_cast<std::size_t>(nnz));
  if (!i.ok()) return i.status();
  auto v = DeviceBuffer(resource, sizeof(E) * static_cast<std::size_t>(nnz));
  if (!v.ok()) return v.status();
  const Status s = fill(o->data(), i->data(), v->data());
  if (!s.ok()) return s;
  return Access::MakeCompressed<E, F>({rows, cols}, nnz, width, std::move(*o), std::move(*i),
                                      std::move(*v));
}

}  // namespace internal_sparse_hip

template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
template <typename E>
Result<HipNarrowCompressedView<E, Format>> HipCompressedArray<Element, Format>::Narrow() const {
  if (!valid()) return Status(ErrorCode::kInvalidState, "A moved-from HIP sparse array has no view");
  if (width_ != HipIndexWidth::k32)
    return Status(ErrorCode::kInvalidState, "A 64-bit-index HIP clone is viewed through view()");
  return internal_sparse_hip::Access::MakeNarrow<E, Format>(
      static_cast<const nnz_t*>(outer_offsets_.data()),
      static_cast<const std::int32_t*>(inner_indices_.data()),
      static_cast<E*>(const_cast<void*>(values_.data())), shape_, nonzeros_);
}

// SpMV algorithm choice (ASC_HIP_ALG_*): kAuto uses the context's plan for the
// matrix (autotuned, learned, or an L2/LDS column-panel analysis; never the
// stream-staged copy of the values) and merge-path otherwise; kOrdered is the
// serial reference's summation, bit for bit; kColumnPanel runs the matrix's
// column-panel analysis, the stream-staged one included.
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

// ---- clones (CudaCloneCsr, providers/cuda.h:1040-1064) ----------------------

// Host (or device) CSR -> a device CSR clone with int64 indices, as the
// reference's CudaCloneCsr; HipIndexWidth::k32 narrows the indices.
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipCsrClone<std::remove_const_t<SourceElement>>> HipCloneCsr(
    SparseHipContext& context, CsrView<SourceElement> source, MemoryResource& resource,
    HipIndexWidth width = HipIndexWidth::k64) {
  using E = std::remove_const_t<SourceElement>;
  const asc_hip_sparse_desc src = internal_sparse_hip::Describe(source);
  auto array = internal_sparse_hip::FillCompressed<E, SparseCompressedFormat::kCsr>(
      resource, source.rows(), source.columns(), source.nnz(), width,
      [&](void* o, void* i, void* v) {
        asc_hip_sparse_desc out{};
        return internal_sparse_hip::ToStatus(asc_sparse_hip_clone_compressed(
            context.native(), &src, static_cast<std::uint8_t>(width), o, i, v, &out));
      });
  if (!array.ok()) return array.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCsrClone<E>{std::move(*array), std::move(*done)};
}

// Host canonical COO -> device.
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipCoordinateClone<std::remove_const_t<SourceElement>>> HipCloneCoordinate(
    SparseHipContext& context, CoordinateView<SourceElement, 2> source, MemoryResource& resource) {
  using E = std::remove_const_t<SourceElement>;
  auto c = internal_sparse_hip::DeviceBuffer(resource, 16 * static_cast<std::size_t>(source.nnz()));
  if (!c.ok()) return c.status();
  auto v = internal_sparse_hip::DeviceBuffer(resource, sizeof(E) * static_cast<std::size_t>(source.nnz()));
  if (!v.ok()) return v.status();
  const asc_hip_sparse_desc src = internal_sparse_hip::Describe(source);
  asc_hip_sparse_desc out{};
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_clone_coordinate(context.native(), &src, c->data(), v->data(), &out));
  if (!s.ok()) return s;
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCoordinateClone<E>{
      internal_sparse_hip::Access::MakeCoordinate<E>(source.shape(), source.nnz(), std::move(*c),
                                                     std::move(*v)),
      std::move(*done)};
}

// ---- compat SpMV y = alpha*A*x + beta*y (blas.h:2031-2081) -----------------

namespace internal_sparse_hip {

inline Result<std::size_t> CsrSpmvWorkspaceSize(SparseHipContext& context,
                                                const asc_hip_sparse_desc& a,
                                                const asc_hip_vector_desc& x,
                                                const asc_hip_vector_desc& y, std::uint8_t alg) {
  const asc_hip_spmv_options opt{alg, ASC_HIP_ACC_PLAIN, {}};
  std::size_t bytes = 0;
  const Status s = ToStatus(
      asc_sparse_hip_csr_spmv_workspace_size(context.native(), &a, &x, &y, &opt, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

inline Result<CompletionEvent> CsrSpmv(SparseHipContext& context, double alpha,
                                       const asc_hip_sparse_desc& a, const asc_hip_vector_desc& x,
                                       double beta, const asc_hip_vector_desc& y,
                                       MutableMemoryView workspace, std::uint8_t alg) {
  const asc_hip_spmv_options opt{alg, ASC_HIP_ACC_PLAIN, {}};
  return Enqueued(context, asc_sparse_hip_csr_spmv(context.native(), alpha, &a, &x, beta, &y,
                                                   workspace.data(), workspace.size(), &opt));
}

}  // namespace internal_sparse_hip

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipCsrSpmvWorkspaceSize(
    SparseHipContext& context, CsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, HipStridedVectorView<Element> output,
    HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  return internal_sparse_hip::CsrSpmvWorkspaceSize(
      context, internal_sparse_hip::Describe(matrix), internal_sparse_hip::Describe(input),
      internal_sparse_hip::Describe(output), static_cast<std::uint8_t>(algorithm));
}
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipCsrSpmvWorkspaceSize(
    SparseHipContext& context, HipNarrowCsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, HipStridedVectorView<Element> output,
    HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  return internal_sparse_hip::CsrSpmvWorkspaceSize(
      context, internal_sparse_hip::Describe(matrix), internal_sparse_hip::Describe(input),
      internal_sparse_hip::Describe(output), static_cast<std::uint8_t>(algorithm));
}

// CudaCsrSpmv (providers/cuda.h:1786-1802).
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipCsrSpmv(
    SparseHipContext& context, Element alpha, CsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, Element beta, HipStridedVectorView<Element> output,
    MutableMemoryView workspace, HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  return internal_sparse_hip::CsrSpmv(context, static_cast<double>(alpha),
                                      internal_sparse_hip::Describe(matrix),
                                      internal_sparse_hip::Describe(input), static_cast<double>(beta),
                                      internal_sparse_hip::Describe(output), workspace,
                                      static_cast<std::uint8_t>(algorithm));
}
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipCsrSpmv(
    SparseHipContext& context, Element alpha, HipNarrowCsrView<const Element> matrix,
    HipStridedVectorView<const Element> input, Element beta, HipStridedVectorView<Element> output,
    MutableMemoryView workspace, HipSpmvAlgorithm algorithm = HipSpmvAlgorithm::kAuto) {
  return internal_sparse_hip::CsrSpmv(context, static_cast<double>(alpha),
                                      internal_sparse_hip::Describe(matrix),
                                      internal_sparse_hip::Describe(input), static_cast<double>(beta),
                                      internal_sparse_hip::Describe(output), workspace,
                                      static_cast<std::uint8_t>(algorithm));
}

// Column-panel analysis of a device CSR matrix for kColumnPanel.  The
// analysis memory must outlive the plan.  With stream staging the plan holds
// a copy of the values: after changing them in place call
// HipCsrSpmvRefreshValues before the next kColumnPanel SpMV (kAuto never
// runs the copy).
template <typename Matrix>
[[nodiscard]] Result<std::size_t> HipCsrSpmvAnalysisSize(SparseHipContext& context,
                                                         const Matrix& matrix,
                                                         const asc_hip_panel_options* options = nullptr) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_csr_spmv_analysis_size(context.native(), &a, options, &bytes));
  if (!s.ok()) return s;
  return bytes;
}
template <typename Matrix>
[[nodiscard]] Status HipCsrSpmvAnalyze(SparseHipContext& context, const Matrix& matrix,
                                       MutableMemoryView analysis,
                                       const asc_hip_panel_options* options = nullptr) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  return internal_sparse_hip::ToStatus(asc_sparse_hip_csr_spmv_analyze(
      context.native(), &a, options, analysis.data(), analysis.size()));
}
template <typename Matrix>
[[nodiscard]] Result<CompletionEvent> HipCsrSpmvRefreshValues(SparseHipContext& context,
                                                              const Matrix& matrix) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  return internal_sparse_hip::Enqueued(context,
                                       asc_sparse_hip_csr_spmv_refresh_values(context.native(), &a));
}

// ---- standardized SpMV y += alpha*op(A)*x (blas.h:1607-1667) ----------------

// CudaSpmv (providers/cuda.h:1616-1646).  CSR with kNone and CSC with a
// transpose need no workspace; the other two transpose into `workspace`
// (size from HipSpmvWorkspaceSize).
template <typename Element, typename Matrix>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipSpmvWorkspaceSize(SparseHipContext& context,
                                                       SparseBlasTranspose transpose,
                                                       const Matrix& matrix,
                                                       SparseBlasVectorView<const Element> input,
                                                       SparseBlasVectorView<Element> output) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_standard_spmv_workspace_size(
      context.native(), internal_sparse_hip::Op(transpose), &a, &x, &y, nullptr, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSpmv(
    SparseHipContext& context, SparseBlasTranspose transpose, Element alpha,
    CompressedSparseView<const Element, Format> matrix, SparseBlasVectorView<const Element> input,
    SparseBlasVectorView<Element> output,
    MutableMemoryView workspace = internal_sparse_hip::NoWorkspace()) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  const asc_hip_vector_desc x = internal_sparse_hip::Describe(input);
  const asc_hip_vector_desc y = internal_sparse_hip::Describe(output);
  const asc_hip_scalar s{static_cast<double>(alpha), 0.0};
  return internal_sparse_hip::Enqueued(
      context, asc_sparse_hip_standard_spmv(context.native(), internal_sparse_hip::Op(transpose), s,
                                            &a, &x, &y, workspace.data(), workspace.size(), nullptr));
}

// ---- standardized SpMM C += alpha*op(A)*B (blas.h:1692-1750) ----------------

// kPlain: register accumulation with fused multiply-adds, within the plain
// bound of the serial result (row-major B and C with k <= 32, k and both
// leading dimensions multiples of 4, 4-element aligned; other shapes run
// kOrdered); kOrdered: the serial reference's per-entry updates, bit for bit.
enum class HipSpmmForm : std::uint8_t {
  kPlain = ASC_HIP_ALG_AUTO,
  kOrdered = ASC_HIP_ALG_ORDERED,
  // The plain form over the matrix's stream-staged column-panel analysis
  // (HipCsrSpmvAnalyze with ASC_HIP_STAGING_STREAM): CSR, op N and the plain
  // kernel's operands only.
  kPanel = ASC_HIP_ALG_PANEL,
};

template <typename Element, typename Matrix>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipSpmmWorkspaceSize(SparseHipContext& context,
                                                       SparseBlasTranspose transpose,
                                                       const Matrix& matrix,
                                                       SparseBlasMatrixView<const Element> input,
                                                       SparseBlasMatrixView<Element> output,
                                                       HipSpmmForm form = HipSpmmForm::kPlain) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  const asc_hip_matrix_desc b = internal_sparse_hip::Describe(input);
  const asc_hip_matrix_desc c = internal_sparse_hip::Describe(output);
  const asc_hip_spmv_options opt{static_cast<std::uint8_t>(form), ASC_HIP_ACC_PLAIN, {}};
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_standard_spmm_workspace_size(
      context.native(), internal_sparse_hip::Op(transpose), &a, &b, &c, &opt, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

// CudaSpmm (providers/cuda.h:1648-1658).  The workspace is optional: with
// less than HipSpmmWorkspaceSize reports, the plain form runs in one pass.
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSpmm(
    SparseHipContext& context, SparseBlasTranspose transpose, Element alpha,
    CompressedSparseView<const Element, Format> matrix, SparseBlasMatrixView<const Element> input,
    SparseBlasMatrixView<Element> output,
    MutableMemoryView workspace = internal_sparse_hip::NoWorkspace(),
    HipSpmmForm form = HipSpmmForm::kPlain) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(matrix);
  const asc_hip_matrix_desc b = internal_sparse_hip::Describe(input);
  const asc_hip_matrix_desc c = internal_sparse_hip::Describe(output);
  const asc_hip_scalar s{static_cast<double>(alpha), 0.0};
  const asc_hip_spmv_options opt{static_cast<std::uint8_t>(form), ASC_HIP_ACC_PLAIN, {}};
  return internal_sparse_hip::Enqueued(
      context, asc_sparse_hip_standard_spmm(context.native(), internal_sparse_hip::Op(transpose), s,
                                            &a, &b, &c, workspace.data(), workspace.size(), &opt));
}

// ---- conversions (compressed.h:1675-1837) ----------------------------------

template <typename Matrix>
[[nodiscard]] Result<std::size_t> HipTransposeWorkspaceSize(SparseHipContext& context,
                                                            const Matrix& source) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(source);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_convert_compressed_workspace_size(context.native(), &a, &bytes));
  if (!s.ok()) return s;
  return bytes;
}

namespace internal_sparse_hip {

template <typename E, SparseCompressedFormat To>
Result<HipCompressedArray<E, To>> Transpose(SparseHipContext& context, const asc_hip_sparse_desc& a,
                                            MemoryResource& resource, MutableMemoryView workspace) {
  return FillCompressed<E, To>(resource, a.rows, a.cols, a.nnz,
                               static_cast<HipIndexWidth>(a.index_bytes),
                               [&](void* o, void* i, void* v) {
                                 asc_hip_sparse_desc d{};
                                 return ToStatus(asc_sparse_hip_convert_compressed(
                                     context.native(), &a, o, i, v, &d, workspace.data(),
                                     workspace.size()));
                               });
}

}  // namespace internal_sparse_hip

// CSR -> CSC (ConvertToCsc(ctx, CsrView), compressed.h:1753-1760) and CSC ->
// CSR (:1727-1734): a stable device transpose, bit-exact with the reference;
// the result keeps the source's index width.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCscClone<Element>> HipConvertToCsc(SparseHipContext& context,
                                                           CsrView<const Element> source,
                                                           MemoryResource& resource,
                                                           MutableMemoryView workspace) {
  auto out = internal_sparse_hip::Transpose<Element, SparseCompressedFormat::kCsc>(
      context, internal_sparse_hip::Describe(source), resource, workspace);
  if (!out.ok()) return out.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCscClone<Element>{std::move(*out), std::move(*done)};
}
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCscClone<Element>> HipConvertToCsc(SparseHipContext& context,
                                                           HipNarrowCsrView<const Element> source,
                                                           MemoryResource& resource,
                                                           MutableMemoryView workspace) {
  auto out = internal_sparse_hip::Transpose<Element, SparseCompressedFormat::kCsc>(
      context, internal_sparse_hip::Describe(source), resource, workspace);
  if (!out.ok()) return out.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCscClone<Element>{std::move(*out), std::move(*done)};
}
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCsrClone<Element>> HipConvertToCsr(SparseHipContext& context,
                                                           CscView<const Element> source,
                                                           MemoryResource& resource,
                                                           MutableMemoryView workspace) {
  auto out = internal_sparse_hip::Transpose<Element, SparseCompressedFormat::kCsr>(
      context, internal_sparse_hip::Describe(source), resource, workspace);
  if (!out.ok()) return out.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCsrClone<Element>{std::move(*out), std::move(*done)};
}

// Canonical device COO -> CSR / CSC (ConvertToCsr/ConvertToCsc(ctx,
// CoordinateView), compressed.h:1675-1708).
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipConvertCoordinateWorkspaceSize(
    SparseHipContext& context, CoordinateView<const Element, 2> source,
    SparseCompressedFormat format, HipIndexWidth width = HipIndexWidth::k64) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(source);
  std::size_t bytes = 0;
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_coordinate_workspace_size(
      context.native(), &a, format == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC,
      static_cast<std::uint8_t>(width), &bytes));
  if (!s.ok()) return s;
  return bytes;
}

namespace internal_sparse_hip {

template <typename E, SparseCompressedFormat To>
Result<HipCompressedArray<E, To>> FromCoordinates(SparseHipContext& context,
                                                  const CoordinateView<const E, 2>& source,
                                                  MemoryResource& resource,
                                                  MutableMemoryView workspace, HipIndexWidth width) {
  const asc_hip_sparse_desc a = Describe(source);
  return FillCompressed<E, To>(
      resource, source.shape()[0], source.shape()[1], source.nnz(), width,
      [&](void* o, void* i, void* v) {
        asc_hip_sparse_desc d{};
        return ToStatus(asc_sparse_hip_convert_coordinate(
            context.native(), &a, To == SparseCompressedFormat::kCsr ? ASC_HIP_CSR : ASC_HIP_CSC,
            static_cast<std::uint8_t>(width), o, i, v, &d, workspace.data(), workspace.size()));
      });
}

}  // namespace internal_sparse_hip

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCsrClone<Element>> HipConvertToCsr(
    SparseHipContext& context, CoordinateView<const Element, 2> source, MemoryResource& resource,
    MutableMemoryView workspace, HipIndexWidth width = HipIndexWidth::k64) {
  auto out = internal_sparse_hip::FromCoordinates<Element, SparseCompressedFormat::kCsr>(
      context, source, resource, workspace, width);
  if (!out.ok()) return out.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCsrClone<Element>{std::move(*out), std::move(*done)};
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCscClone<Element>> HipConvertToCsc(
    SparseHipContext& context, CoordinateView<const Element, 2> source, MemoryResource& resource,
    MutableMemoryView workspace, HipIndexWidth width = HipIndexWidth::k64) {
  auto out = internal_sparse_hip::FromCoordinates<Element, SparseCompressedFormat::kCsc>(
      context, source, resource, workspace, width);
  if (!out.ok()) return out.status();
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCscClone<Element>{std::move(*out), std::move(*done)};
}

// Device CSR/CSC -> canonical COO (ConvertToCoordinate, compressed.h:1781-1837).
template <typename Element, SparseCompressedFormat Format>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<std::size_t> HipConvertToCoordinateWorkspaceSize(
    SparseHipContext& context, CompressedSparseView<const Element, Format> source) {
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(source);
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
    MemoryResource& resource, MutableMemoryView workspace = internal_sparse_hip::NoWorkspace()) {
  auto c = internal_sparse_hip::DeviceBuffer(resource, 16 * static_cast<std::size_t>(source.nnz()));
  if (!c.ok()) return c.status();
  auto v = internal_sparse_hip::DeviceBuffer(resource, sizeof(Element) * static_cast<std::size_t>(source.nnz()));
  if (!v.ok()) return v.status();
  const asc_hip_sparse_desc a = internal_sparse_hip::Describe(source);
  asc_hip_sparse_desc d{};
  const Status s = internal_sparse_hip::ToStatus(asc_sparse_hip_convert_to_coordinate(
      context.native(), &a, c->data(), v->data(), &d, workspace.data(), workspace.size()));
  if (!s.ok()) return s;
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipCoordinateClone<Element>{
      internal_sparse_hip::Access::MakeCoordinate<Element>(source.shape(), source.nnz(),
                                                           std::move(*c), std::move(*v)),
      std::move(*done)};
}

// Device CoordinateBuilder::Finalize (coordinate.h:1300-1497): raw device
// triplets (interleaved int64 (row, col) in insertion order, and values) to a
// canonical coordinate array, with the builder's DuplicatePolicy and
// ExplicitZeroPolicy.  Waits for the device (the entry count is known only
// after the sort); a workspace comes from `resource`.
template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<HipCoordinateArray<Element>> HipFinalizeCoordinates(
    SparseHipContext& context, const index_t* device_coordinates, const Element* device_values,
    extent_t rows, extent_t cols, nnz_t count, DuplicatePolicy duplicates, ExplicitZeroPolicy zeros,
    MemoryResource& resource) {
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
  auto ws = internal_sparse_hip::DeviceBuffer(resource, bytes);
  if (!ws.ok()) return ws.status();
  auto c = internal_sparse_hip::DeviceBuffer(resource, 16 * static_cast<std::size_t>(count));
  if (!c.ok()) return c.status();
  auto v = internal_sparse_hip::DeviceBuffer(resource, sizeof(Element) * static_cast<std::size_t>(count));
  if (!v.ok()) return v.status();
  asc_hip_sparse_desc d{};
  s = internal_sparse_hip::ToStatus(asc_sparse_hip_coo_finalize(
      context.native(), &t,
      duplicates == DuplicatePolicy::kSum ? ASC_HIP_DUPLICATE_SUM : ASC_HIP_DUPLICATE_REJECT,
      zeros == ExplicitZeroPolicy::kDrop ? ASC_HIP_ZERO_DROP : ASC_HIP_ZERO_KEEP, c->data(),
      v->data(), &d, ws->data(), ws->size()));
  if (!s.ok()) return s;
  return internal_sparse_hip::Access::MakeCoordinate<Element>({rows, cols}, d.nnz, std::move(*c),
                                                              std::move(*v));
}

// ---- sparse Level-1 (blas.h:1330-1582; CudaSparseDot ... CudaSparseScatter,
// providers/cuda.h:1464-1614) ----------------------------------------------

// Host canonical indexed vector -> device (CudaCloneIndexedVector,
// providers/cuda.h:1247-1272).
template <typename SourceElement>
  requires internal_sparse_hip::kSupportedElement<SourceElement>
[[nodiscard]] Result<HipIndexedVectorClone<std::remove_const_t<SourceElement>>>
HipCloneIndexedVector(SparseHipContext& context, SparseBlasIndexedVectorView<SourceElement> source,
                      MemoryResource& resource) {
  using E = std::remove_const_t<SourceElement>;
  auto i = internal_sparse_hip::DeviceBuffer(resource, 8 * static_cast<std::size_t>(source.nnz()));
  if (!i.ok()) return i.status();
  auto v = internal_sparse_hip::DeviceBuffer(resource, sizeof(E) * static_cast<std::size_t>(source.nnz()));
  if (!v.ok()) return v.status();
  const asc_hip_indexed_desc src = internal_sparse_hip::Describe(source);
  asc_hip_indexed_desc out{};
  const Status s = internal_sparse_hip::ToStatus(
      asc_sparse_hip_clone_indexed(context.native(), &src, i->data(), v->data(), &out));
  if (!s.ok()) return s;
  auto done = RecordHipEvent(context.execution_context());
  if (!done.ok()) return done.status();
  return HipIndexedVectorClone<E>{internal_sparse_hip::Access::MakeIndexed<E>(
                                      source.dense_extent(), source.nnz(), std::move(*i),
                                      std::move(*v)),
                                  std::move(*done)};
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
                                              internal_sparse_hip::Describe(sparse),
                                              internal_sparse_hip::Describe(dense), &r, algorithm);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseAxpy(SparseHipContext& context, Element alpha,
                                                    SparseBlasIndexedVectorView<const Element> sparse,
                                                    SparseBlasVectorView<Element> dense) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_AXPY, SparseBlasConjugation::kUnconjugated, alpha,
      internal_sparse_hip::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseGather(SparseHipContext& context,
                                                      SparseBlasVectorView<const Element> dense,
                                                      SparseBlasIndexedVectorView<Element> sparse) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_GATHER, SparseBlasConjugation::kUnconjugated, Element(1),
      internal_sparse_hip::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseGatherZero(SparseHipContext& context,
                                                          SparseBlasVectorView<Element> dense,
                                                          SparseBlasIndexedVectorView<Element> sparse) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_GATHER_ZERO, SparseBlasConjugation::kUnconjugated, Element(1),
      internal_sparse_hip::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
      HipSpmvAlgorithm::kAuto);
}

template <typename Element>
  requires internal_sparse_hip::kSupportedElement<Element>
[[nodiscard]] Result<CompletionEvent> HipSparseScatter(SparseHipContext& context,
                                                       SparseBlasIndexedVectorView<const Element> sparse,
                                                       SparseBlasVectorView<Element> dense) {
  return internal_sparse_hip::Level1<Element>(
      context, ASC_HIP_LEVEL1_SCATTER, SparseBlasConjugation::kUnconjugated, Element(1),
      internal_sparse_hip::Describe(sparse), internal_sparse_hip::Describe(dense), nullptr,
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
                                           SparseBlasTranspose transpose, MemoryResource& resource) {
    const asc_hip_sparse_desc desc = internal_sparse_hip::Describe(matrix);
    const std::uint8_t op = internal_sparse_hip::Op(transpose);
    std::size_t bytes = 0;
    Status s = internal_sparse_hip::ToStatus(
        asc_sparse_hip_triangular_analysis_size(context.native(), op, &desc, &bytes));
    if (!s.ok()) return s;
    auto buf = internal_sparse_hip::DeviceBuffer(resource, bytes);
    if (!buf.ok()) return buf.status();
    asc_hip_triangular_info info{};
    s = internal_sparse_hip::ToStatus(asc_sparse_hip_triangular_analyze(
        context.native(), op, triangle == SparseBlasTriangle::kLower ? ASC_HIP_LOWER : ASC_HIP_UPPER,
        diagonal == SparseBlasDiagonal::kUnit ? ASC_HIP_UNIT : ASC_HIP_NON_UNIT, &desc, buf->data(),
        buf->size(), &info));
    if (!s.ok()) return s;
    return HipTriangularPlan(desc, info, op, std::move(*buf));
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
    asc_hip_matrix_desc b{};
    b.data = right_hand_side.data();
    b.rows = right_hand_side.size();
    b.cols = 1;
    b.leading_dimension = right_hand_side.increment();
    b.layout = ASC_HIP_ROW_MAJOR;
    b.element_kind = internal_sparse_hip::ElementKind<Element>();
    b.memory_space = static_cast<std::uint8_t>(right_hand_side.memory_space());
    b.writable = 1;
    const asc_hip_scalar a{static_cast<double>(alpha), 0.0};
    return internal_sparse_hip::Enqueued(
        context, asc_sparse_hip_triangular_solve(context.native(), transpose_, a, &desc_, &b,
                                                 analysis_.data(), analysis_.size(), &info_));
  }

 private:
  HipTriangularPlan(asc_hip_sparse_desc desc, asc_hip_triangular_info info, std::uint8_t op,
                    Buffer analysis) noexcept
      : desc_(desc), info_(info), transpose_(op), analysis_(std::move(analysis)) {}
  asc_hip_sparse_desc desc_;
  asc_hip_triangular_info info_;
  std::uint8_t transpose_;
  Buffer analysis_;
};

}  // namespace asc
