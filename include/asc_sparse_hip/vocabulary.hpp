// Source-compatible declarations of the asc:: vocabulary that the HIP sparse
// facet (asc_sparse_hip/hip.hpp) uses, for builds outside the asc-cpp tree.
// Names, numeric values and member functions follow the reference headers
// (include/asc/core/status.h, types.h, memory.h; include/asc/sparse/
// compressed.h, coordinate.h, blas.h) so that code written against the
// reference's CUDA facet reads the same against this one.  Inside the
// reference tree these declarations are replaced by its own headers.
#pragma once

#include <array>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <span>
#include <string>
#include <type_traits>
#include <utility>

namespace asc {

// include/asc/core/types.h:35-62
using extent_t = std::int64_t;
using stride_t = std::int64_t;
using index_t = std::int64_t;
using nnz_t = std::int64_t;

// include/asc/core/status.h:35-55
enum class ErrorCode : std::int32_t {
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
  kInternal = 18,
};

// include/asc/core/status.h:79-196: code, message, provider, native code.
class Status {
 public:
  Status() = default;
  Status(ErrorCode code, std::string message, std::string provider = {},
         std::int64_t native_code = 0)
      : code_(code), message_(std::move(message)), provider_(std::move(provider)),
        native_code_(native_code) {}
  static Status Ok() { return Status(); }
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

// A value or the Status that explains its absence.
template <typename T>
class Result {
 public:
  Result(T value) : value_(std::move(value)) {}  // NOLINT: implicit like the reference
  Result(Status status) : status_(std::move(status)) {
    if (status_.ok()) status_ = Status(ErrorCode::kInternal, "OK status without a value");
  }
  [[nodiscard]] bool ok() const noexcept { return value_.has_value(); }
  [[nodiscard]] const Status& status() const noexcept { return status_; }
  [[nodiscard]] T& value() & { return *value_; }
  [[nodiscard]] const T& value() const& { return *value_; }
  [[nodiscard]] T&& value() && { return std::move(*value_); }
  T& operator*() & { return *value_; }
  const T& operator*() const& { return *value_; }
  T&& operator*() && { return std::move(*value_); }
  T* operator->() { return &*value_; }
  const T* operator->() const { return &*value_; }

 private:
  std::optional<T> value_;
  Status status_;
};

// include/asc/core/memory.h:30-35
enum class MemorySpace : std::uint8_t {
  kHost = 0,
  kPinnedHost = 1,
  kDevice = 2,
  kManaged = 3,
};

// Byte ranges with their placement (memory.h ConstMemoryView/MutableMemoryView).
struct ConstMemoryView {
  const void* data = nullptr;
  std::size_t size = 0;
  MemorySpace memory_space = MemorySpace::kHost;
  [[nodiscard]] bool valid() const noexcept { return data != nullptr || size == 0; }
};
struct MutableMemoryView {
  void* data = nullptr;
  std::size_t size = 0;
  MemorySpace memory_space = MemorySpace::kDevice;
  [[nodiscard]] bool valid() const noexcept { return data != nullptr || size == 0; }
};

// include/asc/sparse/coordinate.h:86-89
enum class SparseCompressedFormat : std::uint8_t { kCsr, kCsc };

// include/asc/sparse/blas.h:88-106
enum class SparseBlasTranspose : std::uint8_t { kNone, kTranspose, kConjugateTranspose };
enum class SparseBlasLayout : std::uint8_t { kColumnMajor, kRowMajor };

namespace internal_hip_vocabulary {

// Canonical structure of a host compressed operand (compressed.h:711-746):
// offsets start at 0, never decrease and end at nnz; inner indices are in
// range and strictly increase within each segment.
inline Status ValidateCompressed(const nnz_t* offsets, const index_t* indices, extent_t outer,
                                 extent_t inner, nnz_t nnz) {
  if (offsets[0] != 0 || offsets[outer] != nnz)
    return Status(ErrorCode::kInvalidArgument, "Compressed offsets must start at 0 and end at nnz");
  for (extent_t o = 0; o < outer; ++o) {
    if (offsets[o + 1] < offsets[o])
      return Status(ErrorCode::kInvalidArgument, "Compressed offsets must not decrease");
    for (nnz_t p = offsets[o]; p < offsets[o + 1]; ++p) {
      if (indices[p] < 0 || indices[p] >= inner)
        return Status(ErrorCode::kIndex, "Compressed inner index out of range");
      if (p > offsets[o] && indices[p] <= indices[p - 1])
        return Status(ErrorCode::kInvalidArgument,
                      "Compressed inner indices must increase within a segment");
    }
  }
  return Status();
}

// Value-pointer replacement checks shared by the compressed and coordinate
// views (compressed.h:472-516, coordinate.h:533-569): a non-wrapping span,
// no overlap with the structure spans, and no partial overlap with the
// current values.
inline bool SpansOverlap(const void* a, std::size_t an, const void* b, std::size_t bn) noexcept {
  if (an == 0 || bn == 0) return false;
  const auto x = reinterpret_cast<std::uintptr_t>(a), y = reinterpret_cast<std::uintptr_t>(b);
  if (an > UINTPTR_MAX - x || bn > UINTPTR_MAX - y) return true;
  return x < y + bn && y < x + an;
}

inline Status CheckRebind(const void* values, const void* current, std::size_t value_bytes,
                          const void* first, std::size_t first_bytes, const void* second,
                          std::size_t second_bytes, const char* structure_message) {
  if (value_bytes == 0) return Status();
  if (values == nullptr)
    return Status(ErrorCode::kInvalidArgument, "Nonempty sparse storage cannot be null");
  if (value_bytes > UINTPTR_MAX - reinterpret_cast<std::uintptr_t>(values))
    return Status(ErrorCode::kOverflow, "Sparse storage address range overflows uintptr_t");
  if (SpansOverlap(first, first_bytes, values, value_bytes) ||
      SpansOverlap(second, second_bytes, values, value_bytes))
    return Status(ErrorCode::kInvalidArgument, structure_message);
  if (values != current && SpansOverlap(current, value_bytes, values, value_bytes))
    return Status(ErrorCode::kInvalidArgument,
                  "Rebound sparse values cannot partially overlap existing values");
  return Status();
}

}  // namespace internal_hip_vocabulary

class HipProviderAccess;

// Non-owning canonical CSR/CSC view (compressed.h:83-755).  Host views are
// validated by Create; device views are trusted only when a provider made
// them (compressed.h:759-770).
template <typename Element, SparseCompressedFormat Format>
class CompressedSparseView {
 public:
  static Result<CompressedSparseView> Create(const nnz_t* outer_offsets,
                                             const index_t* inner_indices, Element* values,
                                             std::span<const extent_t, 2> shape, nnz_t nonzeros,
                                             MemorySpace memory_space) {
    if (shape[0] < 0 || shape[1] < 0)
      return Status(ErrorCode::kShape, "Sparse shape extents must be nonnegative");
    if (nonzeros < 0 || static_cast<__int128>(shape[0]) * shape[1] < nonzeros)
      return Status(ErrorCode::kShape, "Sparse NNZ is incompatible with compressed shape");
    CompressedSparseView v;
    v.offsets_ = outer_offsets;
    v.indices_ = inner_indices;
    v.values_ = values;
    v.shape_ = {shape[0], shape[1]};
    v.nonzeros_ = nonzeros;
    v.memory_space_ = memory_space;
    if (memory_space == MemorySpace::kHost || memory_space == MemorySpace::kPinnedHost) {
      if (outer_offsets == nullptr || (nonzeros > 0 && (inner_indices == nullptr || values == nullptr)))
        return Status(ErrorCode::kInvalidArgument, "Compressed storage pointers are null");
      Status s = internal_hip_vocabulary::ValidateCompressed(outer_offsets, inner_indices,
                                                             v.outer(), v.inner(), nonzeros);
      if (!s.ok()) return s;
    }
    return v;
  }

  [[nodiscard]] constexpr extent_t rows() const noexcept { return shape_[0]; }
  [[nodiscard]] constexpr extent_t columns() const noexcept { return shape_[1]; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr const nnz_t* outer_offsets() const noexcept { return offsets_; }
  [[nodiscard]] constexpr const index_t* inner_indices() const noexcept { return indices_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }
  [[nodiscard]] constexpr bool canonical_structure_trusted() const noexcept { return trusted_; }
  [[nodiscard]] constexpr extent_t outer() const noexcept {
    return Format == SparseCompressedFormat::kCsr ? shape_[0] : shape_[1];
  }
  [[nodiscard]] constexpr extent_t inner() const noexcept {
    return Format == SparseCompressedFormat::kCsr ? shape_[1] : shape_[0];
  }
  // A read-only view of the same storage.
  operator CompressedSparseView<const Element, Format>() const noexcept {  // NOLINT
    CompressedSparseView<const Element, Format> v;
    v.offsets_ = offsets_;
    v.indices_ = indices_;
    v.values_ = values_;
    v.shape_ = shape_;
    v.nonzeros_ = nonzeros_;
    v.memory_space_ = memory_space_;
    v.trusted_ = trusted_;
    v.index_bytes_ = index_bytes_;
    return v;
  }
  // Inner-index width of the stored indices (8, or 4 after a narrowing clone).
  [[nodiscard]] constexpr int index_bytes() const noexcept { return index_bytes_; }

  // The same structure, shape, NNZ, memory space and trust over other values
  // (compressed.h:472-516); neither allocates nor copies.
  template <typename ReboundElement>
    requires std::same_as<std::remove_const_t<ReboundElement>, std::remove_const_t<Element>>
  [[nodiscard]] Result<CompressedSparseView<ReboundElement, Format>> RebindValues(
      ReboundElement* values) const {
    const std::size_t n = static_cast<std::size_t>(nonzeros_);
    if (n > SIZE_MAX / sizeof(Element))
      return Status(ErrorCode::kOverflow, "Sparse value byte count overflows");
    Status s = internal_hip_vocabulary::CheckRebind(
        values, values_, n * sizeof(Element), offsets_,
        static_cast<std::size_t>(outer() + 1) * sizeof(nnz_t), indices_,
        n * static_cast<std::size_t>(index_bytes_),
        "Rebound sparse values cannot overlap compressed structure");
    if (!s.ok()) return s;
    CompressedSparseView<ReboundElement, Format> v;
    v.offsets_ = offsets_;
    v.indices_ = indices_;
    v.values_ = values;
    v.shape_ = shape_;
    v.nonzeros_ = nonzeros_;
    v.memory_space_ = memory_space_;
    v.trusted_ = trusted_;
    v.index_bytes_ = index_bytes_;
    return v;
  }

 private:
  template <typename, SparseCompressedFormat>
  friend class CompressedSparseView;
  friend class HipProviderAccess;
  const nnz_t* offsets_ = nullptr;
  const index_t* indices_ = nullptr;
  Element* values_ = nullptr;
  std::array<extent_t, 2> shape_{0, 0};
  nnz_t nonzeros_ = 0;
  MemorySpace memory_space_ = MemorySpace::kHost;
  bool trusted_ = false;
  int index_bytes_ = 8;
};

template <typename Element>
using CsrView = CompressedSparseView<Element, SparseCompressedFormat::kCsr>;
template <typename Element>
using CscView = CompressedSparseView<Element, SparseCompressedFormat::kCsc>;

// Canonical rank-2 coordinate view (coordinate.h:266-790): interleaved
// (row, col) pairs in strictly increasing lexicographic order.
template <typename Element, std::size_t Rank>
class CoordinateView {
  static_assert(Rank == 2, "the HIP facet handles matrices");

 public:
  static Result<CoordinateView> Create(const index_t* coordinates, Element* values,
                                       std::span<const extent_t, Rank> extents, nnz_t nonzeros,
                                       MemorySpace memory_space) {
    if (extents[0] < 0 || extents[1] < 0 || nonzeros < 0)
      return Status(ErrorCode::kShape, "Coordinate extents and NNZ must be nonnegative");
    if (memory_space == MemorySpace::kHost || memory_space == MemorySpace::kPinnedHost) {
      for (nnz_t k = 0; k < nonzeros; ++k) {
        const index_t r = coordinates[2 * k], c = coordinates[2 * k + 1];
        if (r < 0 || r >= extents[0] || c < 0 || c >= extents[1])
          return Status(ErrorCode::kIndex, "Coordinate out of range");
        if (k > 0) {
          const index_t pr = coordinates[2 * k - 2], pc = coordinates[2 * k - 1];
          if (r < pr || (r == pr && c <= pc))
            return Status(ErrorCode::kInvalidArgument,
                          "Coordinates must be strictly increasing (canonical order)");
        }
      }
    }
    CoordinateView v;
    v.coordinates_ = coordinates;
    v.values_ = values;
    v.extents_ = {extents[0], extents[1]};
    v.nonzeros_ = nonzeros;
    v.memory_space_ = memory_space;
    return v;
  }
  [[nodiscard]] constexpr const index_t* coordinates() const noexcept { return coordinates_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] constexpr extent_t extent(std::size_t d) const noexcept { return extents_[d]; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }
  [[nodiscard]] constexpr bool canonical_structure_trusted() const noexcept { return trusted_; }

  // The same coordinates, extents, NNZ, memory space and trust over other
  // values (coordinate.h:533-569).
  template <typename ReboundElement>
    requires std::same_as<std::remove_const_t<ReboundElement>, std::remove_const_t<Element>>
  [[nodiscard]] Result<CoordinateView<ReboundElement, Rank>> RebindValues(
      ReboundElement* values) const {
    const std::size_t n = static_cast<std::size_t>(nonzeros_);
    if (n > SIZE_MAX / (Rank * sizeof(index_t)))
      return Status(ErrorCode::kOverflow, "Sparse coordinate byte count overflows");
    Status s = internal_hip_vocabulary::CheckRebind(
        values, values_, n * sizeof(Element), coordinates_, n * Rank * sizeof(index_t), nullptr,
        0, "Rebound sparse values cannot overlap coordinates");
    if (!s.ok()) return s;
    CoordinateView<ReboundElement, Rank> v;
    v.coordinates_ = coordinates_;
    v.values_ = values;
    v.extents_ = extents_;
    v.nonzeros_ = nonzeros_;
    v.memory_space_ = memory_space_;
    v.trusted_ = trusted_;
    return v;
  }

 private:
  template <typename, std::size_t>
  friend class CoordinateView;
  friend class HipProviderAccess;
  const index_t* coordinates_ = nullptr;
  Element* values_ = nullptr;
  std::array<extent_t, 2> extents_{0, 0};
  nnz_t nonzeros_ = 0;
  MemorySpace memory_space_ = MemorySpace::kHost;
  bool trusted_ = false;
};

// Strided vector with its backing storage (blas.h:276-515).  A negative
// increment addresses element i at logical_first + i*increment.
template <typename Element>
class SparseBlasVectorView {
 public:
  static Result<SparseBlasVectorView> Create(Element* logical_first, extent_t size,
                                             stride_t increment, ConstMemoryView backing_storage) {
    if (size < 0 || increment == 0 || !backing_storage.valid())
      return Status(ErrorCode::kInvalidArgument, "A sparse BLAS vector descriptor is invalid");
    SparseBlasVectorView v;
    v.data_ = logical_first;
    v.size_ = size;
    v.increment_ = increment;
    v.memory_space_ = backing_storage.memory_space;
    return v;
  }
  [[nodiscard]] constexpr Element* data() const noexcept { return data_; }
  [[nodiscard]] constexpr extent_t size() const noexcept { return size_; }
  [[nodiscard]] constexpr stride_t increment() const noexcept { return increment_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }

 private:
  Element* data_ = nullptr;
  extent_t size_ = 0;
  stride_t increment_ = 1;
  MemorySpace memory_space_ = MemorySpace::kHost;
};

// include/asc/sparse/blas.h:74-77, 117-134
enum class SparseBlasConjugation : std::uint8_t { kUnconjugated, kConjugated };
enum class SparseBlasTriangle : std::uint8_t { kUpper, kLower };
enum class SparseBlasDiagonal : std::uint8_t { kNonUnit, kUnit };

// Canonical indexed (sparse) vector (blas.h:526-789): nnz strictly
// increasing indices in [0, dense_extent) and their values.  Host views are
// checked by Create; device views are trusted only when a provider made them.
template <typename Element>
class SparseBlasIndexedVectorView {
 public:
  static Result<SparseBlasIndexedVectorView> Create(const index_t* indices, Element* values,
                                                    nnz_t nonzeros, extent_t dense_extent,
                                                    ConstMemoryView index_storage,
                                                    ConstMemoryView value_storage) {
    if (nonzeros < 0 || dense_extent < 0 || nonzeros > dense_extent ||
        index_storage.memory_space != value_storage.memory_space)
      return Status(ErrorCode::kInvalidArgument, "A sparse indexed-vector descriptor is invalid");
    SparseBlasIndexedVectorView v;
    v.indices_ = indices;
    v.values_ = values;
    v.nonzeros_ = nonzeros;
    v.dense_extent_ = dense_extent;
    v.memory_space_ = index_storage.memory_space;
    if (v.memory_space_ == MemorySpace::kHost || v.memory_space_ == MemorySpace::kPinnedHost) {
      index_t previous = -1;
      for (nnz_t p = 0; p < nonzeros; ++p) {
        if (indices[p] <= previous || indices[p] < 0 || indices[p] >= dense_extent)
          return Status(ErrorCode::kIndex,
                        "Sparse indexed-vector indices must be sorted, unique, zero-based, and in range");
        previous = indices[p];
      }
    }
    return v;
  }
  [[nodiscard]] constexpr const index_t* indices() const noexcept { return indices_; }
  [[nodiscard]] constexpr Element* values() const noexcept { return values_; }
  [[nodiscard]] constexpr nnz_t nnz() const noexcept { return nonzeros_; }
  [[nodiscard]] constexpr extent_t dense_extent() const noexcept { return dense_extent_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }
  [[nodiscard]] constexpr bool canonical_structure_trusted() const noexcept { return trusted_; }
  operator SparseBlasIndexedVectorView<const Element>() const noexcept {  // NOLINT
    SparseBlasIndexedVectorView<const Element> v;
    v.indices_ = indices_;
    v.values_ = values_;
    v.nonzeros_ = nonzeros_;
    v.dense_extent_ = dense_extent_;
    v.memory_space_ = memory_space_;
    v.trusted_ = trusted_;
    return v;
  }

 private:
  template <typename>
  friend class SparseBlasIndexedVectorView;
  friend class HipProviderAccess;
  const index_t* indices_ = nullptr;
  Element* values_ = nullptr;
  nnz_t nonzeros_ = 0;
  extent_t dense_extent_ = 0;
  MemorySpace memory_space_ = MemorySpace::kHost;
  bool trusted_ = false;
};

// Dense block, row- or column-major with a leading dimension (blas.h:816-1063).
template <typename Element>
class SparseBlasMatrixView {
 public:
  static Result<SparseBlasMatrixView> Create(Element* data, extent_t rows, extent_t columns,
                                             SparseBlasLayout layout, stride_t leading_dimension,
                                             ConstMemoryView backing_storage) {
    const extent_t minor = layout == SparseBlasLayout::kRowMajor ? columns : rows;
    if (rows < 0 || columns < 0 || leading_dimension < (minor > 1 ? minor : 1) ||
        !backing_storage.valid())
      return Status(ErrorCode::kInvalidArgument, "A sparse BLAS matrix descriptor is invalid");
    SparseBlasMatrixView v;
    v.data_ = data;
    v.rows_ = rows;
    v.columns_ = columns;
    v.layout_ = layout;
    v.leading_dimension_ = leading_dimension;
    v.memory_space_ = backing_storage.memory_space;
    return v;
  }
  [[nodiscard]] constexpr Element* data() const noexcept { return data_; }
  [[nodiscard]] constexpr extent_t rows() const noexcept { return rows_; }
  [[nodiscard]] constexpr extent_t columns() const noexcept { return columns_; }
  [[nodiscard]] constexpr SparseBlasLayout layout() const noexcept { return layout_; }
  [[nodiscard]] constexpr stride_t leading_dimension() const noexcept { return leading_dimension_; }
  [[nodiscard]] constexpr MemorySpace memory_space() const noexcept { return memory_space_; }

 private:
  Element* data_ = nullptr;
  extent_t rows_ = 0;
  extent_t columns_ = 0;
  SparseBlasLayout layout_ = SparseBlasLayout::kColumnMajor;
  stride_t leading_dimension_ = 1;
  MemorySpace memory_space_ = MemorySpace::kHost;
};

}  // namespace asc
