// CSR SpMV kernels for gfx950.
//
//   vector<G>  G lanes (1..64) per row, coalesced strided walk of the row,
//              butterfly reduction, fused alpha/beta epilogue.
//   merge      merge-path split (Merrill & Garland, SC'16): every tile owns
//              exactly kTile items of the rows+nnz merge sequence, so row
//              length skew cannot unbalance the grid.  Products are staged
//              in LDS, each thread walks kIpt consecutive merge items, a
//              wave64 segmented scan hands partial rows to their owners and
//              rows that cross tiles are finished by a small fix-up kernel.
//   ordered    one lane per row, strict left-to-right unfused chain: the
//              serial reference's arithmetic bit for bit
//              (src/sparse/reference_blas.cc:13-39, blas.h:1653-1665).
//   panel      column-panel plan: one vector<G> pass per column range of x,
//              from a split table built once per matrix, so every pass
//              gathers from an x panel resident in L2.
//
// Value storage V in {double, float, bf16, f16}; accumulation in double for
// double and in float otherwise (plain, or compensated with TwoSum plus the
// FMA product error).
#include <vector>

#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;
// Row blocks per panel in one panel-kernel launch (the grid stays well below
// the 2^32 work-items a dispatch can address; blocks stride beyond it).
constexpr int64_t kPanelRowBlockCap = int64_t(1) << 20;

// ---- accumulators ---------------------------------------------------------
template <typename A>
struct PlainAcc {
  A s;  // trivially constructible so it can live in LDS; start from zero()
  __device__ __forceinline__ static PlainAcc zero() {
    PlainAcc r;
    r.s = A(0);
    return r;
  }
  __device__ __forceinline__ void add(A p) { s += p; }
  __device__ __forceinline__ void add_product(A v, A x) { s += v * x; }
  __device__ __forceinline__ void merge_after(const PlainAcc& later) {
    s += later.s;
  }
  __device__ __forceinline__ PlainAcc shfl_xor(int off, int width) const {
    PlainAcc r;
    r.s = __shfl_xor(s, off, width);
    return r;
  }
  __device__ __forceinline__ PlainAcc shfl_up(int off) const {
    PlainAcc r;
    r.s = __shfl_up(s, off);
    return r;
  }
  __device__ __forceinline__ double total_f64() const {
    return static_cast<double>(s);
  }
  __device__ __forceinline__ A total() const { return s; }
  __device__ __forceinline__ static PlainAcc from_f64(double v) {
    PlainAcc r;
    r.s = static_cast<A>(v);
    return r;
  }
};

// Compensated FP32 accumulation: s + c carries the running sum; every
// product's rounding error is captured with an FMA and every addition's with
// Knuth's TwoSum.
struct TwoSumAcc {
  float s;
  float c;
  __device__ __forceinline__ static TwoSumAcc zero() {
    TwoSumAcc r;
    r.s = 0.0f;
    r.c = 0.0f;
    return r;
  }
  __device__ __forceinline__ static void two_sum(float a, float b, float& t,
                                                 float& e) {
#pragma clang fp contract(off)
    t = a + b;
    const float bp = t - a;
    e = (a - (t - bp)) + (b - bp);
  }
  __device__ __forceinline__ void add(float p) {
#pragma clang fp contract(off)
    float t, e;
    two_sum(s, p, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void add_pair(float p, float pe) {
#pragma clang fp contract(off)
    float t, e;
    two_sum(s, p, t, e);
    s = t;
    c += e + pe;
  }
  __device__ __forceinline__ void add_product(float v, float x) {
#pragma clang fp contract(off)
    const float p = v * x;
    const float pe = __builtin_fmaf(v, x, -p);
    add_pair(p, pe);
  }
  __device__ __forceinline__ void merge_after(const TwoSumAcc& later) {
#pragma clang fp contract(off)
    float t, e;
    two_sum(s, later.s, t, e);
    s = t;
    c += e + later.c;
  }
  __device__ __forceinline__ TwoSumAcc shfl_xor(int off, int width) const {
    TwoSumAcc r;
    r.s = __shfl_xor(s, off, width);
    r.c = __shfl_xor(c, off, width);
    return r;
  }
  __device__ __forceinline__ TwoSumAcc shfl_up(int off) const {
    TwoSumAcc r;
    r.s = __shfl_up(s, off);
    r.c = __shfl_up(c, off);
    return r;
  }
  __device__ __forceinline__ double total_f64() const {
    return static_cast<double>(s) + static_cast<double>(c);
  }
  __device__ __forceinline__ float total() const { return s + c; }
  // (s, c) with s + c == v to within FP32's rounding of the tail.
  __device__ __forceinline__ static TwoSumAcc from_f64(double v) {
    TwoSumAcc r;
    r.s = static_cast<float>(v);
    r.c = static_cast<float>(v - static_cast<double>(r.s));
    return r;
  }
};

template <typename V, bool TWOSUM>
struct AccSel {
  using type = PlainAcc<typename acc_of<V>::type>;
};
template <typename V>
struct AccSel<V, true> {
  using type = TwoSumAcc;
};

struct KParams {
  const int64_t* offsets;
  const void* indices;
  const void* values;
  int64_t rows;
  int64_t nnz;
  const void* x;
  int64_t incx;
  void* y;
  int64_t incy;
  double alpha;
  double beta;
  int form;  // SpmvForm
  // 16- or 32-bit values on a 2-value boundary and int32 indices on an
  // 8-byte boundary: the vector kernel loads entries in pairs.
  int paired;
};

// Fused epilogue for one finished row.
//   compat:   y = alpha*sum (+ beta*y when beta != 0)
//   standard: y = y + alpha*sum
// FP64 storage computes in FP64; FP32-accumulated forms compute in FP32, and
// the compensated form finishes its (s, c) pair in FP64 (one multiply-add
// per row: the only FP64 arithmetic of the reduced-precision path).
template <typename V, typename Acc>
__device__ __forceinline__ void finish_row(const KParams& p, int64_t row,
                                           const Acc& acc) {
  using A = typename acc_of<V>::type;
  A* y = static_cast<A*>(p.y) + row * p.incy;
  if constexpr (std::is_same<Acc, TwoSumAcc>::value) {
    const double s = acc.total_f64();
    double r;
    if (p.form == 0) {
      r = p.alpha * s;
      if (p.beta != 0.0) r += p.beta * static_cast<double>(*y);
    } else {
      r = static_cast<double>(*y) + p.alpha * s;
    }
    *y = static_cast<float>(r);
  } else {
    const A alpha = static_cast<A>(p.alpha);
    const A beta = static_cast<A>(p.beta);
    const A s = acc.total();
    if (p.form == 0) {
      A r = alpha * s;
      if (beta != A(0)) r += beta * *y;
      *y = r;
    } else {
      *y = *y + alpha * s;
    }
  }
}

// Adds a carried partial (rows split across merge tiles) to a finished y.
template <typename V, typename Acc>
__device__ __forceinline__ void add_carry(const KParams& p, int64_t row,
                                          const Acc& acc) {
  using A = typename acc_of<V>::type;
  A* y = static_cast<A*>(p.y) + row * p.incy;
  if constexpr (std::is_same<Acc, TwoSumAcc>::value) {
    *y = static_cast<float>(static_cast<double>(*y) + p.alpha * acc.total_f64());
  } else {
    *y = *y + static_cast<A>(p.alpha) * acc.total();
  }
}

// ---------------------------------------------------------------------------
// vector<G>: G lanes per row.
// ---------------------------------------------------------------------------
// The G-lane sum of entries [b, e), reduced to every lane of the group.  Two
// independent chains keep two gathers in flight per lane.
template <int G, typename V, typename I, typename Acc>
__device__ __forceinline__ Acc group_sum(const KParams& p, int64_t b, int64_t e, int lane) {
  using A = typename acc_of<V>::type;
  const V* __restrict__ val = static_cast<const V*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const A* __restrict__ x = static_cast<const A*>(p.x);
  constexpr bool kNt = G >= 16;
  Acc acc = Acc::zero();
  Acc acc2 = Acc::zero();
  int64_t k = b + lane;
  for (; k + G < e; k += 2 * G) {
    const int64_t c0 = load_index_as<kNt>(col, k);
    const int64_t c1 = load_index_as<kNt>(col, k + G);
    const A v0 = widen<A>(ld_matrix<kNt>(val + k));
    const A v1 = widen<A>(ld_matrix<kNt>(val + k + G));
    acc.add_product(v0, x[c0 * p.incx]);
    acc2.add_product(v1, x[c1 * p.incx]);
  }
  if (k < e) {
    const int64_t c0 = load_index_as<kNt>(col, k);
    acc.add_product(widen<A>(ld_matrix<kNt>(val + k)), x[c0 * p.incx]);
  }
  acc.merge_after(acc2);
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc.merge_after(acc.shfl_xor(off, G));
  return acc;
}

// 16- or 32-bit values with int32 indices (p.paired): lane l takes the entry
// pairs (2l, 2l+1) from the row's first even entry, one load for the two
// values and one 8-byte load for the two indices, half the load instructions
// of the per-entry loop; an odd first or last entry goes to lane 0.  Loads
// follow the per-entry loop's policy: nontemporal from 16 lanes per row up,
// cached below (where neighbouring rows share the lines).
template <int G, typename Acc, typename V>
__device__ __forceinline__ Acc group_sum_pairs(const KParams& p, int64_t b, int64_t e, int lane) {
  using A = typename acc_of<V>::type;
  static_assert(sizeof(V) == 2 || sizeof(V) == 4, "pairs of 16- or 32-bit values");
  typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
  typedef int32_t i32x2_t __attribute__((ext_vector_type(2)));
  const V* __restrict__ val = static_cast<const V*>(p.values);
  const int32_t* __restrict__ col = static_cast<const int32_t*>(p.indices);
  const A* __restrict__ x = static_cast<const A*>(p.x);
  Acc acc = Acc::zero();
  Acc acc2 = Acc::zero();
  int64_t b2 = b;
  if ((b & 1) && b < e) {
    if (lane == 0) acc.add_product(widen<A>(val[b]), x[static_cast<int64_t>(col[b]) * p.incx]);
    b2 = b + 1;
  }
  constexpr bool kNt = G >= 16;
  for (int64_t k = b2 + 2 * lane; k + 1 < e; k += 2 * G) {
    const i32x2_t cc = ld_matrix<kNt>(reinterpret_cast<const i32x2_t*>(col + k));
    V v0, v1;
    if constexpr (sizeof(V) == 2) {
      const uint32_t w = ld_matrix<kNt>(reinterpret_cast<const uint32_t*>(val + k));
      v0.bits = static_cast<uint16_t>(w & 0xFFFFu);
      v1.bits = static_cast<uint16_t>(w >> 16);
    } else {
      const u32x2_t w = ld_matrix<kNt>(reinterpret_cast<const u32x2_t*>(val + k));
      v0 = __uint_as_float(w.x);
      v1 = __uint_as_float(w.y);
    }
    acc.add_product(widen<A>(v0), x[static_cast<int64_t>(cc.x) * p.incx]);
    acc2.add_product(widen<A>(v1), x[static_cast<int64_t>(cc.y) * p.incx]);
  }
  if (((e - b2) & 1) && e > b2 && lane == 0)
    acc.add_product(widen<A>(val[e - 1]), x[static_cast<int64_t>(col[e - 1]) * p.incx]);
  acc.merge_after(acc2);
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc.merge_after(acc.shfl_xor(off, G));
  return acc;
}

// Rows of at least `skip_from` entries are left to the panel kernels
// (INT64_MAX: none are).
template <int G, typename V, typename I, bool TWOSUM>
__global__ __launch_bounds__(kBlock) void spmv_vector_kernel(KParams p, int64_t skip_from) {
  using Acc = typename AccSel<V, TWOSUM>::type;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = blk * (kBlock / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (row >= p.rows) return;  // whole G-groups leave together
  const int64_t b = ld_matrix<(G >= 16)>(p.offsets + row);
  const int64_t e = ld_matrix<(G >= 16)>(p.offsets + row + 1);
  if (e - b >= skip_from) return;
  Acc acc = Acc::zero();
  // Pairs halve the lanes a row keeps busy: only up to 16 lanes per row.
  if constexpr (sizeof(V) <= 4 && sizeof(I) == 4 && G <= 16) {
    if (p.paired) acc = group_sum_pairs<G, Acc, V>(p, b, e, lane);  // uniform branch
    else acc = group_sum<G, V, I, Acc>(p, b, e, lane);
  } else {
    acc = group_sum<G, V, I, Acc>(p, b, e, lane);
  }
  if (lane == 0) finish_row<V>(p, row, acc);
}

// ---------------------------------------------------------------------------
// Column-panel plan for long rows.  Random columns of a large x miss in L2 on
// nearly every gather.  The columns are cut into `panels` ranges (a multiple
// of 8) and the rows with at least `long_min` entries are listed; each listed
// row's entries in one panel form a segment, found from the split table.
// Each (panel, row) partial sum goes to an FP64 buffer; the combine kernel
// adds them in panel order and applies the row epilogue.  Shorter rows (a
// small share of the entries of a skewed matrix) run through vector<G> in one
// pass.  Two forms walk the segments:
//   L2:  block b works on panel 8*group + b % 8 for one block of listed rows.
//        Work-groups are dispatched to the 8 XCDs round robin, so each XCD
//        gathers from one panel of x at a time, which stays in its L2, and
//        the 8 XCDs read adjacent slices of the same rows at about the same
//        time.
//   LDS: a work-group stages its panel of x (at most kLdsPanelBytes) in LDS
//        and walks one chunk of listed rows, gathering from LDS.  Work-groups
//        are ordered chunk-major, so every panel of a chunk runs at about the
//        same time (the chunk's matrix lines are fetched from HBM about once)
//        and XCD k stages the panels congruent to k mod 8, which its L2 keeps.
// ---------------------------------------------------------------------------
template <int G, typename V, typename I, bool TWOSUM>
__global__ __launch_bounds__(kBlock) void spmv_panel_kernel(
    KParams p, const int64_t* __restrict__ row_begin, const uint32_t* __restrict__ splits,
    int64_t nlong, int64_t panels, int64_t row_blocks, int64_t rb_grid,
    double* __restrict__ partial) {
  using Acc = typename AccSel<V, TWOSUM>::type;
  const int64_t slot = blockIdx.x % 8;
  const int64_t rest = blockIdx.x / 8;
  const int64_t q = (rest / rb_grid) * 8 + slot;
  if (q >= panels) return;
  const int lane = threadIdx.x % G;
  // rb_grid blocks per panel stride over the row blocks (launch-size cap).
  for (int64_t rbk = rest % rb_grid; rbk < row_blocks; rbk += rb_grid) {
    const int64_t i = rbk * (kBlock / G) + threadIdx.x / G;
    if (i < nlong) {  // uniform over the G-group
      const int64_t rb = row_begin[i];
      const int64_t b = q == 0 ? rb : rb + ld_stream(splits + (q - 1) * nlong + i);
      const int64_t e = rb + ld_stream(splits + q * nlong + i);
      const Acc acc = group_sum<G, V, I, Acc>(p, b, e, lane);
      if (lane == 0) partial[q * nlong + i] = acc.total_f64();
    }
  }
}

constexpr int kLdsBlock = 1024;

// The G-lane sum of entries [b, e) of one panel segment, x read from the LDS
// copy `sx` of columns [col0, col0 + panel width).
template <int G, typename V, typename I, typename Acc, typename A>
__device__ __forceinline__ Acc group_sum_lds(const V* __restrict__ val, const I* __restrict__ col,
                                             const A* sx, int64_t col0, int64_t b, int64_t e,
                                             int lane) {
  Acc acc = Acc::zero();
  Acc acc2 = Acc::zero();
  int64_t k = b + lane;
  for (; k + G < e; k += 2 * G) {
    const int c0 = static_cast<int>(load_index(col, k) - col0);
    const int c1 = static_cast<int>(load_index(col, k + G) - col0);
    const A v0 = widen<A>(ld_stream(val + k));
    const A v1 = widen<A>(ld_stream(val + k + G));
    acc.add_product(v0, sx[c0]);
    acc2.add_product(v1, sx[c1]);
  }
  if (k < e) {
    const int c0 = static_cast<int>(load_index(col, k) - col0);
    acc.add_product(widen<A>(ld_stream(val + k)), sx[c0]);
  }
  acc.merge_after(acc2);
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc.merge_after(acc.shfl_xor(off, G));
  return acc;
}

template <int G, typename V, typename I, bool TWOSUM>
__global__ __launch_bounds__(kLdsBlock) void spmv_panel_lds_kernel(
    KParams p, int64_t cols, const int64_t* __restrict__ row_begin,
    const uint32_t* __restrict__ splits, int64_t nlong, int64_t panels, int64_t panel_cols,
    int64_t chunk_rows, double* __restrict__ partial) {
  using A = typename acc_of<V>::type;
  using Acc = typename AccSel<V, TWOSUM>::type;
  __shared__ A sx[kLdsPanelBytes / sizeof(A)];
  const int64_t q = blockIdx.x % panels;
  const int64_t chunk = blockIdx.x / panels;
  const int64_t col0 = q * panel_cols;
  const int64_t width = cols - col0 < panel_cols ? cols - col0 : panel_cols;
  const A* __restrict__ x = static_cast<const A*>(p.x);
  for (int64_t k = threadIdx.x; k < width; k += kLdsBlock) sx[k] = x[(col0 + k) * p.incx];
  __syncthreads();
  const V* __restrict__ val = static_cast<const V*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const int64_t r0 = chunk * chunk_rows;
  const int64_t r1 = nlong - r0 < chunk_rows ? nlong : r0 + chunk_rows;
  const int lane = threadIdx.x % G;
  constexpr int kRowsPerPass = kLdsBlock / G;
  int64_t i = r0 + threadIdx.x / G;
  // The next pass's segment bounds load while this pass runs.
  int64_t b = 0, e = 0;
  if (i < r1) {
    const int64_t rb = row_begin[i];
    b = q == 0 ? rb : rb + ld_stream(splits + (q - 1) * nlong + i);
    e = rb + ld_stream(splits + q * nlong + i);
  }
  for (; i < r1; i += kRowsPerPass) {  // uniform over the G-group
    const int64_t inext = i + kRowsPerPass;
    int64_t bn = 0, en = 0;
    if (inext < r1) {
      const int64_t rb = row_begin[inext];
      bn = q == 0 ? rb : rb + ld_stream(splits + (q - 1) * nlong + inext);
      en = rb + ld_stream(splits + q * nlong + inext);
    }
    const Acc acc = group_sum_lds<G, V, I, Acc, A>(val, col, sx, col0, b, e, lane);
    if (lane == 0) partial[q * nlong + i] = acc.total_f64();
    b = bn;
    e = en;
  }
}

template <typename V, bool TWOSUM>
__global__ __launch_bounds__(kBlock) void spmv_panel_combine_kernel(
    KParams p, const uint32_t* __restrict__ long_rows, int64_t nlong, int64_t panels,
    const double* __restrict__ partial) {
  using Acc = typename AccSel<V, TWOSUM>::type;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nlong) return;
  double s = 0.0;
  for (int64_t q = 0; q < panels; ++q) s += partial[q * nlong + i];
  finish_row<V>(p, long_rows[i], Acc::from_f64(s));
}

// ---- column-panel analysis ------------------------------------------------
__global__ __launch_bounds__(kBlock) void panel_flag_kernel(
    const int64_t* __restrict__ offsets, int64_t rows, int64_t long_min,
    uint32_t* __restrict__ flags) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (r < rows) flags[r] = offsets[r + 1] - offsets[r] >= long_min ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void panel_list_kernel(
    const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, int64_t rows,
    uint32_t* __restrict__ long_rows, uint32_t* __restrict__ count) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= rows) return;
  if (flags[r]) long_rows[pos[r]] = static_cast<uint32_t>(r);
  if (r == rows - 1) *count = pos[r] + flags[r];
}

// Row starts and split table: one thread per listed row, an advancing binary
// search per panel boundary over the row's sorted columns.
template <typename I>
__global__ __launch_bounds__(kBlock) void panel_split_kernel(
    const int64_t* __restrict__ offsets, const I* __restrict__ col,
    const uint32_t* __restrict__ long_rows, int64_t nlong, int64_t panel_cols,
    int64_t panels, int64_t* __restrict__ row_begin, uint32_t* __restrict__ splits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nlong) return;
  const int64_t row = long_rows[i];
  const int64_t b = offsets[row];
  const int64_t e = offsets[row + 1];
  row_begin[i] = b;
  int64_t lo = b;
  for (int64_t q = 1; q < panels; ++q) {
    const int64_t bound = q * panel_cols;
    int64_t hi = e;
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo) >> 1);
      if (static_cast<int64_t>(col[mid]) < bound)
        lo = mid + 1;
      else
        hi = mid;
    }
    splits[(q - 1) * nlong + i] = static_cast<uint32_t>(lo - b);
  }
  splits[(panels - 1) * nlong + i] = static_cast<uint32_t>(e - b);
}

// ---------------------------------------------------------------------------
// Stream form of the column-panel plan.  The analysis copies the listed rows'
// entries panel-major (every panel's entries contiguous, rows ascending inside
// it), each as a 16-bit column within the panel (bit 15 flags the first entry
// of a (panel, row) segment) and the value.  A work-group stages its panel of
// x in LDS and each of its waves walks one slice of the panel's stream in
// steps of kStreamTile entries, kStreamEpl consecutive entries per lane:
// fully coalesced loads, x gathered from LDS, a wave64 segmented scan over
// the head flags, and every finished segment's sum written to its (panel,
// row) slot.  A segment still open at a slice's end leaves its part in
// carry[slice]; the fix-up kernel adds such parts to the slot, and the
// combine kernel adds the slots of a row in panel order.
// ---------------------------------------------------------------------------
typedef double f64x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

static_assert(kStreamEpl == 8, "the stream loads below move 8 entries per lane");

// Values are stored lane-interleaved inside each step of kStreamTile entries:
// with w = 16 / sizeof(V) values per 16-byte word, entry e = 8*l + w*k + r of
// a step sits at k*64*w + l*w + r, so the k-th 16-byte load of a wave reads
// one contiguous kiB and still hands lane l its own eight consecutive entries.
template <typename V>
__host__ __device__ __forceinline__ int64_t stream_value_pos(int64_t pos) {
  constexpr int64_t w = 16 / static_cast<int64_t>(sizeof(V));
  const int64_t base = pos & ~static_cast<int64_t>(kStreamTile - 1);
  const int64_t e = pos & (kStreamTile - 1);
  const int64_t l = e / kStreamEpl, j = e % kStreamEpl;
  return base + (j / w) * kWave * w + l * w + j % w;
}

__device__ __forceinline__ void stream_values(const double* p, int lane, double (&v)[kStreamEpl]) {
  const f64x2_t* q = reinterpret_cast<const f64x2_t*>(p) + lane;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const f64x2_t a = __builtin_nontemporal_load(q + k * kWave);
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}
__device__ __forceinline__ void stream_values(const float* p, int lane, float (&v)[kStreamEpl]) {
  const f32x4_t* q = reinterpret_cast<const f32x4_t*>(p) + lane;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const f32x4_t a = __builtin_nontemporal_load(q + k * kWave);
    v[4 * k] = a.x;
    v[4 * k + 1] = a.y;
    v[4 * k + 2] = a.z;
    v[4 * k + 3] = a.w;
  }
}
// 16-bit storage: eight values in one 16-byte load, two per 32-bit word.
__device__ __forceinline__ u32x4_t stream_words(const void* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
}
__device__ __forceinline__ uint32_t word_of(const u32x4_t& w, int k) {
  return k == 0 ? w.x : k == 1 ? w.y : k == 2 ? w.z : w.w;
}
__device__ __forceinline__ void stream_values(const bf16_t* p, int lane, float (&v)[kStreamEpl]) {
  const u32x4_t w = stream_words(p + kStreamEpl * lane);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(word_of(w, k) << 16);
    v[2 * k + 1] = __uint_as_float(word_of(w, k) & 0xFFFF0000u);
  }
}
__device__ __forceinline__ void stream_values(const f16_t* p, int lane, float (&v)[kStreamEpl]) {
  const u32x4_t w = stream_words(p + kStreamEpl * lane);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f16_t h;
    h.bits = static_cast<uint16_t>(word_of(w, k) & 0xFFFFu);
    v[2 * k] = to_float(h);
    h.bits = static_cast<uint16_t>(word_of(w, k) >> 16);
    v[2 * k + 1] = to_float(h);
  }
}

template <typename V>
struct StreamStep {
  u32x4_t c;  // eight 16-bit columns
  typename acc_of<V>::type v[kStreamEpl];
};

// The step starting at entry t (a multiple of kStreamTile from the panel's
// start): lane's eight columns and values.
// The empty asm with a memory clobber keeps the compiler from sinking the
// loads down to their first use (which would leave nothing in flight while
// the previous step is reduced); it emits no instruction.
template <typename V>
__device__ __forceinline__ void stream_load(const uint16_t* __restrict__ scol,
                                            const V* __restrict__ sval, int64_t t, int lane,
                                            StreamStep<V>& s) {
  s.c = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(scol + t + kStreamEpl * lane));
  stream_values(sval + t, lane, s.v);
  asm volatile("" ::: "memory");
}

// DPP lane moves (VALU, no LDS traffic): the value of the source lane the
// control selects, or `old` where the source lies outside the row / wave or
// the row is masked off.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_move(float v, float old) {
  return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v), __float_as_int(old)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v, double old) {
  const int64_t b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = dpp_move<CTRL, ROWS>(static_cast<int>(b), static_cast<int>(o));
  const int hi = dpp_move<CTRL, ROWS>(static_cast<int>(b >> 32), static_cast<int>(o >> 32));
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) |
                              static_cast<int64_t>(static_cast<uint32_t>(lo)));
}
// gfx9 DPP controls: row_shr:n, row_bcast:15, row_bcast:31, wave_shr:1, wave_shl:1.
constexpr int kDppRowShr = 0x110;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

// Branch-free segmented scan.  Each lane carries its sum s and a pass-through
// factor m (1 while no head is at or before it in the scanned range, else 0),
// kept as a value whose high word is the only nonzero part (1.0 or 0.0).  A
// step folds in the left neighbour's pair: s += m * s_left, m = m * m_left.
// Row shifts read zeros past the row start (bound control); the row
// broadcasts leave masked-off rows with the identity (0, 1).
constexpr uint32_t kOneHi64 = 0x3FF00000u;  // high word of 1.0 (FP64)
constexpr uint32_t kOne32 = 0x3F800000u;    // 1.0f

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64_or_zero(double v) {
  const int64_t b = __double_as_longlong(v);
  int lo, hi;
  if constexpr (ROWS == 0xF) {
    lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, true);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROWS, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROWS, 0xF, false);
  }
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) |
                              static_cast<int64_t>(static_cast<uint32_t>(lo)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f64_or_zero(float v) {  // FP32 sums
  if constexpr (ROWS == 0xF)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
  else
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

template <typename A> struct PassMask;
template <> struct PassMask<double> {
  static constexpr uint32_t kOne = kOneHi64;
  __device__ static double value(uint32_t w) {
    return __longlong_as_double(static_cast<int64_t>(w) << 32);
  }
};
template <> struct PassMask<float> {
  static constexpr uint32_t kOne = kOne32;
  __device__ static float value(uint32_t w) { return __uint_as_float(w); }
};

template <int CTRL, int ROWS, typename A>
__device__ __forceinline__ void pass_scan_step(A& s, uint32_t& m) {
  const A ls = dpp_f64_or_zero<CTRL, ROWS>(s);
  const uint32_t lm = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(PassMask<A>::kOne), static_cast<int>(m), CTRL, ROWS, 0xF, false));
  s = __builtin_fma(ls, PassMask<A>::value(m), s);
  m &= lm;
}
template <typename A>
__device__ __forceinline__ void pass_scan_wave(A& s, uint32_t& m) {
  pass_scan_step<kDppRowShr | 1, 0xF>(s, m);
  pass_scan_step<kDppRowShr | 2, 0xF>(s, m);
  pass_scan_step<kDppRowShr | 4, 0xF>(s, m);
  pass_scan_step<kDppRowShr | 8, 0xF>(s, m);
  pass_scan_step<kDppRowBcast15, 0xA>(s, m);
  pass_scan_step<kDppRowBcast31, 0xC>(s, m);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) |
                              static_cast<int64_t>(static_cast<uint32_t>(lo)));
}
__device__ __forceinline__ float read_lane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Waits for a step's loads to land.  Stores and loads share one counter on
// this target, so a wait issued after the next step's loads would also wait
// for those: the current step's data is claimed first, then the next step's
// loads go out and stay in flight while the current step is reduced.
template <typename V>
__device__ __forceinline__ void stream_claim(const StreamStep<V>& s) {
  asm volatile("" ::"v"(s.c.x), "v"(s.c.y), "v"(s.c.z), "v"(s.c.w));
#pragma unroll
  for (int j = 0; j < kStreamEpl; ++j) asm volatile("" ::"v"(s.v[j]));
}

// Inclusive scan of the lanes' head counts (DPP, VALU only).
template <int CTRL, int ROWS>
__device__ __forceinline__ void count_step(int& n) {
  n += dpp_move<CTRL, ROWS>(n, 0);
}
__device__ __forceinline__ void count_wave(int& n) {
  count_step<kDppRowShr | 1, 0xF>(n);
  count_step<kDppRowShr | 2, 0xF>(n);
  count_step<kDppRowShr | 4, 0xF>(n);
  count_step<kDppRowShr | 8, 0xF>(n);
  count_step<kDppRowBcast15, 0xA>(n);
  count_step<kDppRowBcast31, 0xC>(n);
}

// Segment sums accumulate in a per-wave circular LDS ring of kStreamRing
// slots (slot = ordinal mod kStreamRing).  A streaming wave never stores to
// HBM in its loop: stores and loads share one wait counter on this target, and
// a store pending beside the prefetched loads would make every wait for a
// step's data also wait for the steps behind it.  Instead the work-group's
// writer wave moves finished sums from the rings to HBM.  Per streaming wave,
// three LDS words hand ordinals over:
//   pub:  ordinals below it are final in the ring (set by the streaming wave);
//   con:  ordinals below it are written out, their slots free (writer wave);
//   done: the streaming wave has published its last ordinal.
// Live ordinals [con, end] of a wave never span more than a ring.
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ uint32_t ring_slot(uint32_t ordinal) {
  return ordinal & (kStreamRing - 1u);
}
// The writer stores whole 128-byte lines of partial sums (16 ordinals) while
// a streaming wave runs: a line written in two parts at different times can
// reach HBM as two partial-line writes, each a read-modify-write.
constexpr uint32_t kLineOrdinals = 16;

struct RingCtl {
  lds_u32* pub;
  lds_u32* con;
  lds_u32* done;
};

__device__ __forceinline__ uint32_t lds_get(lds_u32* p) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(
      static_cast<int>(*static_cast<volatile lds_u32*>(p))));
}
// Lane 0 writes (one LDS word; all lanes hold the same value).
__device__ __forceinline__ void lds_put(lds_u32* p, uint32_t v, int lane) {
  if (lane == 0) *static_cast<volatile lds_u32*>(p) = v;
}
// Completes this wave's earlier LDS operations before its later ones issue
// (lgkmcnt(0) only: vector-memory loads stay in flight).
__device__ __forceinline__ void lds_complete() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  asm volatile("" ::: "memory");
}
// Completes this wave's vector-memory operations (vmcnt(0)).
__device__ __forceinline__ void vmem_complete() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0x0F70);
  asm volatile("" ::: "memory");
}

// One step: kStreamTile entries from t, kStreamEpl consecutive per lane
// (FULL: all of them inside the slice).  Ordinals come from a scan of the
// lanes' head counts: o0 is the segment open at the lane's start, and the
// lane's k-th head opens o0 + k.  Each lane then walks its entries once:
//   - a segment that opens and closes inside the lane is written to its slot;
//   - the run after the lane's last head is written to that segment's slot
//     (its first contribution);
//   - the run before the first head (or the whole lane, without heads) is
//     added to the slot of o0 with an LDS atomic, after all writes.
// The open segment keeps accumulating in its slot from step to step.  A step
// whose segments would not fit in the ring even once the writer has emptied
// it (more than a ring of segments in 512 entries) writes them to HBM
// directly instead, with a segmented scan of the lanes' runs.
struct StepHeads {
  uint32_t hmask;  // bit j: entry j opens a segment
  int own;         // heads in the lane
  uint32_t o0;     // segment open at the lane's start
  uint32_t end;    // segment open after the step
  bool direct;     // the step writes its segments straight to HBM
  bool gave_up;    // the wait for the writer reached its bound
};

// Records a wave that gave up waiting (any lane; every writer stores the
// same value).
__device__ __forceinline__ void stream_fault(uint32_t* fault) {
  if (fault != nullptr)
    __hip_atomic_fetch_or(fault, kFaultStreamWait, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The step's heads and ordinals.  Publishes the ordinals the previous steps
// finished, and waits for the writer while the step would wrap the ring onto
// ordinals it has not written out yet (con_c caches the writer's progress).
template <bool FULL, typename V>
__device__ __forceinline__ StepHeads stream_heads(const StreamStep<V>& cur, int64_t t,
                                                  int64_t s1, int lane, const RingCtl& ctl,
                                                  uint32_t open, uint32_t& con_c,
                                                  uint32_t spin_limit) {
  const int64_t e = t + kStreamEpl * lane;
  uint32_t hmask = 0;  // bit j: entry j opens a segment
#pragma unroll
  for (int j = 0; j < kStreamEpl; ++j) {
    const uint32_t c = (word_of(cur.c, j / 2) >> (16 * (j & 1))) & 0xFFFFu;
    bool hj = (c & kStreamHead) != 0;
    if constexpr (!FULL) hj = hj && e + j < s1;
    hmask |= hj ? 1u << j : 0u;
  }
  const int own = __builtin_popcount(hmask);
  int incl = own;
  count_wave(incl);
  const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(incl, kWave - 1));
  const uint32_t o0 = open + static_cast<uint32_t>(incl - own);
  // One wave's LDS operations take effect in issue order, so the ring updates
  // of the previous steps land before this publication (the compiler is kept
  // from moving them past it).
  asm volatile("" ::: "memory");
  lds_put(ctl.pub, open, lane);
  const uint32_t end = open + total;
  // Slots in use run from the oldest ordinal the writer still owns (con; a
  // slice that opens on a head starts with con one past open) or the open
  // segment, whichever is older, to end.  With pub at open the writer always
  // catches up to open's line (it keeps back less than a line); a step that
  // still does not fit goes direct.
  const auto span = [&](uint32_t c) {
    return end - (static_cast<int32_t>(c - open) < 0 ? c : open);
  };
  bool direct = false, gave_up = false;
  if (span(con_c) >= static_cast<uint32_t>(kStreamRing)) {
    const uint32_t line = open & ~(kLineOrdinals - 1u);
    // con_c and the count are wave-uniform (readfirstlane), so the whole
    // wave leaves the loop together.
    for (uint32_t spins = 0;; ++spins) {
      con_c = lds_get(ctl.con);
      if (span(con_c) < static_cast<uint32_t>(kStreamRing) ||
          static_cast<int32_t>(con_c - line) >= 0)
        break;
      if (spins >= spin_limit) {
        gave_up = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    direct = span(con_c) >= static_cast<uint32_t>(kStreamRing);
  }
  StepHeads hd;
  hd.hmask = hmask;
  hd.own = own;
  hd.o0 = o0;
  hd.end = end;
  hd.direct = __builtin_amdgcn_readfirstlane(direct ? 1 : 0) != 0;  // wave-uniform
  hd.gave_up = __builtin_amdgcn_readfirstlane(gave_up ? 1 : 0) != 0;
  return hd;
}

template <bool FULL, typename V, typename A>
__device__ __forceinline__ void stream_values(const StreamStep<V>& cur, const StepHeads& hd,
                                              const A* sx, int64_t t, int64_t s0, int64_t s1,
                                              int lane, lds_f64* ring, const RingCtl& ctl,
                                              double* __restrict__ partial, uint32_t& open,
                                              uint32_t& con_c) {
  const int64_t e = t + kStreamEpl * lane;
  const uint32_t hmask = hd.hmask;
  const int own = hd.own;
  const uint32_t o0 = hd.o0;
  const uint32_t end = hd.end;
  const bool h0 = (hmask & 1u) != 0;
  const uint32_t spare = kStreamRing + lane;
  A s = A(0), lead = A(0);
  uint32_t seen = 0;
  if (!hd.direct) {
#pragma unroll
    for (int j = 0; j < kStreamEpl; ++j) {
      const uint32_t c = (word_of(cur.c, j / 2) >> (16 * (j & 1))) & 0x7FFFu;
      A prod = cur.v[j] * sx[c];
      if constexpr (!FULL) prod = e + j < s1 ? prod : A(0);
      const bool hj = (hmask >> j) & 1u;
      ring[hj && seen > 0 ? ring_slot(o0 + seen) : spare] = static_cast<double>(s);
      lead = hj && seen == 0 ? s : lead;
      s = __builtin_fma(s, PassMask<A>::value(hj ? 0u : PassMask<A>::kOne), prod);
      seen += hj ? 1u : 0u;
    }
    ring[own > 0 ? ring_slot(o0 + static_cast<uint32_t>(own)) : spare] = static_cast<double>(s);
    __hip_atomic_fetch_add(&ring[ring_slot(o0)], static_cast<double>(own > 0 ? lead : s),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  } else {
    // The writer has written all it will of this wave's ring (up to con, the
    // line of open): this wave writes the rest of [con, open) and the step's
    // ordinals itself, and hands the ring back at end.
    const int32_t kept = static_cast<int32_t>(open - con_c);
    for (int32_t k = lane; k < kept; k += kWave)
      partial[con_c + static_cast<uint32_t>(k)] =
          static_cast<double>(ring[ring_slot(con_c + static_cast<uint32_t>(k))]);
    const A carry_in = static_cast<A>(ring[ring_slot(open)]);
#pragma unroll
    for (int j = 0; j < kStreamEpl; ++j) {
      const uint32_t c = (word_of(cur.c, j / 2) >> (16 * (j & 1))) & 0x7FFFu;
      A prod = cur.v[j] * sx[c];
      if constexpr (!FULL) prod = e + j < s1 ? prod : A(0);
      const bool hj = (hmask >> j) & 1u;
      if (hj && seen > 0) partial[o0 + seen] = static_cast<double>(s);
      lead = hj && seen == 0 ? s : lead;
      s = __builtin_fma(s, PassMask<A>::value(hj ? 0u : PassMask<A>::kOne), prod);
      seen += hj ? 1u : 0u;
    }
    // Segmented scan of the trailing runs: what reaches each lane from the left.
    A sc = s;
    uint32_t pass = own > 0 ? 0u : PassMask<A>::kOne;
    pass_scan_wave(sc, pass);
    sc = __builtin_fma(carry_in, PassMask<A>::value(pass), sc);
    const A from_left = dpp_move<kDppWaveShr1, 0xF>(sc, carry_in);  // lane 0: the open segment
    const int next_h0 = dpp_move<kDppWaveShl1, 0xF>(h0 ? 1 : 0, 0);
    // The run before the first head closes there (a head on entry 0 closes
    // the segment the lane to the left carried, written by that lane, or for
    // lane 0 the open segment of the previous step; a slice that opens on a
    // head leaves that segment to the previous slice).
    if (own > 0 && (!h0 || (lane == 0 && t > s0)))
      partial[o0] = static_cast<double>(from_left + lead);
    // The trailing run closes when the next lane opens on a head.
    if (next_h0 != 0) partial[o0 + own] = static_cast<double>(own > 0 ? s : from_left + s);
    const A open_sum = read_lane(sc, kWave - 1);
    if (lane == 0) ring[ring_slot(end)] = static_cast<double>(open_sum);
    // These stores complete here, so no store is pending at the next step's
    // wait for its data.
    vmem_complete();
    lds_complete();
    lds_put(ctl.con, end, lane);  // before pub: the writer reads pub first
    lds_put(ctl.pub, end, lane);
    con_c = end;
  }
  open = end;
}

// The writer wave: moves each streaming wave's published ordinals from its
// ring to HBM as contiguous bursts until every streaming wave is done.
__device__ __forceinline__ void stream_writer(int waves, int lane, const double* sring,
                                              lds_u32* pub, lds_u32* con, lds_u32* done,
                                              double* __restrict__ partial,
                                              uint32_t spin_limit, uint32_t* fault) {
  uint32_t finished = 0;
  const uint32_t all = (1u << waves) - 1u;
  const int w_of = lane < kStreamWaves ? lane : 0;
  if (spin_limit == 0) return;  // test setting: no writer
  // Passes in a row that found nothing to write: bounded like the streaming
  // waves' waits (a streaming wave always reaches done, or gives up).
  uint32_t idle = 0;
  while (finished != all) {
    // Lane w reads wave w's done word, then (after it returns) pub and con:
    // pub is final once done was seen set.
    const uint32_t dv = *static_cast<volatile lds_u32*>(done + w_of);
    lds_complete();
    const uint32_t pv = *static_cast<volatile lds_u32*>(pub + w_of);
    const uint32_t cv = *static_cast<volatile lds_u32*>(con + w_of);
    bool worked = false;
    for (int w = 0; w < waves; ++w) {
      if ((finished >> w) & 1u) continue;
      const uint32_t d = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(dv), w));
      const uint32_t p = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(pv), w));
      const uint32_t c = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cv), w));
      // Whole lines only until the wave is done (pub is then final).
      const uint32_t upto = d != 0 ? p : p & ~(kLineOrdinals - 1u);
      if (static_cast<int32_t>(upto - c) > 0) {
        const lds_f64* ring = (const lds_f64*)(sring + w * (kStreamRing + kWave));
        const uint32_t n = upto - c;
        // Four 64-ordinal chunks per pass: their LDS reads go out together,
        // then their stores (one LDS round trip per 256 ordinals).
        for (uint32_t k0 = 0; k0 < n; k0 += 4 * kWave) {
          double v[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const uint32_t k = k0 + static_cast<uint32_t>(lane + j * kWave);
            v[j] = k < n ? static_cast<double>(ring[ring_slot(c + k)]) : 0.0;
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const uint32_t k = k0 + static_cast<uint32_t>(lane + j * kWave);
            if (k < n) __builtin_nontemporal_store(v[j], partial + static_cast<int64_t>(c) + k);
          }
        }
        lds_complete();  // the slots are read before they are handed back
        lds_put(con + w, upto, lane);
        worked = true;
      }
      if (d != 0) finished |= 1u << w;
    }
    if (worked) {
      idle = 0;
    } else {
      if (++idle > spin_limit) {
        if (lane == 0) stream_fault(fault);
        return;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

template <typename V>
__global__ __launch_bounds__(kStreamBlock) void spmv_stream_kernel(
    const typename acc_of<V>::type* __restrict__ x, int64_t incx, int64_t cols,
    int64_t panel_cols, const uint16_t* __restrict__ scol, const V* __restrict__ sval,
    const uint32_t* __restrict__ slice_info, const uint32_t* __restrict__ wg_table,
    double* __restrict__ partial, double* __restrict__ carry, uint32_t spin_limit,
    uint32_t* __restrict__ fault) {
  using A = typename acc_of<V>::type;
  __shared__ A sx[kStreamPanelBytes / sizeof(A)];
  __shared__ double sring[kStreamWaves * (kStreamRing + kWave)];
  __shared__ uint32_t sctl[3 * kStreamWaves];
  const uint32_t first = wg_table[2 * blockIdx.x];
  const uint32_t count = wg_table[2 * blockIdx.x + 1];
  const int64_t q = slice_info[4 * static_cast<int64_t>(first) + 3] & ~kStreamClosed;
  // Wave-uniform as far as the compiler can see, so the slice bounds and the
  // step loop's exit stay scalar (a divergent exit would merge every step
  // buffer at the loop latch).
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
  const int lane = threadIdx.x % kWave;
  const bool active = wave < static_cast<int>(count);  // the writer wave never is
  const int64_t slice = static_cast<int64_t>(first) + (active ? wave : 0);
  const uint32_t* si = slice_info + 4 * slice;
  const int64_t s0 = si[0], s1 = si[1];
  lds_u32* pub = (lds_u32*)(sctl);  // address-space casts
  lds_u32* con = (lds_u32*)(sctl + kStreamWaves);
  lds_u32* done = (lds_u32*)(sctl + 2 * kStreamWaves);
  // The first steps load while the panel of x is staged.  A wave loads
  // kStreamDepth - 1 steps ahead, clamped to its slice's last step (the
  // repeated loads of that step hit in cache and are never used).
  const int64_t s_last = s0 + (s1 - 1 - s0) / kStreamTile * kStreamTile;
  static_assert(kStreamDepth == 4, "the step rotation below uses four buffers");
  StreamStep<V> b0, b1, b2, b3;
  uint32_t first_col = 0;  // the slice's first entry (its head flag)
  if (active) {
    first_col = scol[s0];
    stream_load(scol, sval, s0, lane, b0);
    stream_load(scol, sval, min(s0 + kStreamTile, s_last), lane, b1);
    stream_load(scol, sval, min(s0 + 2 * kStreamTile, s_last), lane, b2);
    lds_put(pub + wave, si[2], lane);
    lds_put(con + wave, si[2], lane);
    lds_put(done + wave, 0u, lane);
  }
  const int64_t col0 = q * panel_cols;
  const int64_t width = cols - col0 < panel_cols ? cols - col0 : panel_cols;
  const A* __restrict__ xp = x + col0 * incx;
  int64_t k = threadIdx.x;
  for (; k + 3 * kStreamBlock < width; k += 4 * kStreamBlock) {
    const A x0 = xp[k * incx];
    const A x1 = xp[(k + kStreamBlock) * incx];
    const A x2 = xp[(k + 2 * kStreamBlock) * incx];
    const A x3 = xp[(k + 3 * kStreamBlock) * incx];
    sx[k] = x0;
    sx[k + kStreamBlock] = x1;
    sx[k + 2 * kStreamBlock] = x2;
    sx[k + 3 * kStreamBlock] = x3;
  }
  for (; k < width; k += kStreamBlock) sx[k] = xp[k * incx];
  __syncthreads();
  if (wave == kStreamWaves) {
    stream_writer(static_cast<int>(count < kStreamWaves ? count : kStreamWaves), lane, sring, pub,
                  con, done, partial, spin_limit, fault);
    return;
  }
  if (!active) return;  // whole waves leave together
  const bool closed = (si[3] & kStreamClosed) != 0;
  lds_f64* ring = (lds_f64*)(sring + wave * (kStreamRing + kWave));  // address-space cast
  const RingCtl ctl{pub + wave, con + wave, done + wave};
  uint32_t open = si[2];
  // The segment open at the slice's start gets no first write in this slice;
  // a slice that opens on a head leaves that segment to the previous slice,
  // so the writer starts after it.
  if (lane == 0) ring[ring_slot(open)] = 0.0;
  uint32_t con_c = open + ((first_col & kStreamHead) != 0 ? 1u : 0u);
  lds_put(con + wave, con_c, lane);  // before any publication
  // Full steps, each claiming its buffer and sending the load kStreamDepth-1
  // steps ahead into the buffer the previous step used; the last, partial
  // step of the slice after the loop.
  int64_t t = s0;
  StreamStep<V> last;
#define ASC_STREAM_STEP(CUR, NEXT)                                                          \
  if (t + kStreamTile >= s1) {                                                              \
    last = CUR;                                                                             \
    goto tail;                                                                              \
  }                                                                                         \
  stream_claim(CUR);                                                                        \
  {                                                                                         \
    const StepHeads hd =                                                                    \
        stream_heads<true, V>(CUR, t, s1, lane, ctl, open, con_c, spin_limit);              \
    if (hd.gave_up) goto gave_up;                                                           \
    stream_load(scol, sval, min(t + 3 * kStreamTile, s_last), lane, NEXT);                  \
    stream_values<true, V, A>(CUR, hd, sx, t, s0, s1, lane, ring, ctl, partial, open, con_c); \
  }                                                                                         \
  t += kStreamTile;
  for (;;) {
    ASC_STREAM_STEP(b0, b3)
    ASC_STREAM_STEP(b1, b0)
    ASC_STREAM_STEP(b2, b1)
    ASC_STREAM_STEP(b3, b2)
  }
#undef ASC_STREAM_STEP
tail:
  stream_claim(last);
  {
    const StepHeads hd =
        stream_heads<false, V>(last, t, s1, lane, ctl, open, con_c, spin_limit);
    if (hd.gave_up) goto gave_up;
    stream_values<false, V, A>(last, hd, sx, t, s0, s1, lane, ring, ctl, partial, open, con_c);
  }
  lds_complete();
  lds_put(ctl.pub, open, lane);
  lds_complete();
  lds_put(ctl.done, 1u, lane);
  if (lane == 0) {
    const double v = ring[ring_slot(open)];
    if (closed) partial[open] = v;
    else carry[slice] = v;
  }
  vmem_complete();  // the unused prefetches land before the wave ends
  return;
gave_up:
  // No further ring or HBM writes: record the fault, let the writer finish
  // what was published, and leave once the prefetches have landed.
  if (lane == 0) stream_fault(fault);
  lds_complete();
  lds_put(ctl.done, 1u, lane);
  vmem_complete();
}

// Adds the parts of segments that cross slice ends: the first slice of each
// run of carries into one segment sums the run and adds it to the segment.
__global__ __launch_bounds__(kBlock) void spmv_stream_fixup_kernel(
    const uint32_t* __restrict__ slice_info, int64_t slices, const double* __restrict__ carry,
    double* __restrict__ partial) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= slices) return;
  if (slice_info[4 * k + 3] & kStreamClosed) return;  // an open slice has a successor in its panel
  const uint32_t target = slice_info[4 * (k + 1) + 2];
  if (k > 0 && !(slice_info[4 * (k - 1) + 3] & kStreamClosed) &&
      slice_info[4 * k + 2] == target)
    return;  // not the run's first slice
  double sum = carry[k];
  for (int64_t j = k + 1;
       !(slice_info[4 * j + 3] & kStreamClosed) && slice_info[4 * (j + 1) + 2] == target; ++j)
    sum += carry[j];
  partial[target] += sum;
}

// Row epilogue of the stream form.  Group g is listed rows [g*R, g*R + R)
// (R = kStreamGroupRows).  In each panel the segments of those rows form one
// contiguous ordinal range, ctable[q*(groups+1) + g] .. [q*(groups+1) + g + 1],
// and each row has at most one segment there.  A work-group takes gpb
// adjacent groups (adjacent ranges share cache lines), 16 / gpb waves per
// group; each wave takes a contiguous run of its group's panels (a multiple
// of 8) and adds every segment's sum to its row in an LDS accumulator of its
// own, panel by panel, and the rows then sum their waves' accumulators in
// wave order.  The order is fixed, so the result is the same every run.
template <typename V>
__global__ __launch_bounds__(kLdsBlock) void spmv_stream_combine_kernel(
    KParams p, const uint32_t* __restrict__ long_rows, int64_t nlong, int64_t panels,
    const uint32_t* __restrict__ ctable, int64_t groups, const uint16_t* __restrict__ seg_lrow,
    const double* __restrict__ partial, int gpb) {
  using Acc = typename AccSel<V, false>::type;
  static_assert(kLdsBlock == kCombineWaves * kWave, "one accumulator per combine wave");
  __shared__ double sacc[kCombineWaves * kStreamGroupRows];
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
  const int lane = threadIdx.x % kWave;
  const int wpg = kCombineWaves / gpb;  // waves per group
  const int64_t g = static_cast<int64_t>(blockIdx.x) * gpb + wave / wpg;
  const int part = wave % wpg;
  lds_f64* acc = (lds_f64*)(sacc + wave * kStreamGroupRows);  // address-space cast
  for (int k = lane; k < kStreamGroupRows; k += kWave) acc[k] = 0.0;
  const int64_t stride = groups + 1;
  // This wave's panels [qa, qb): runs of a multiple of 8 (panels is one);
  // none past the last group.
  constexpr int kB = 8;
  const int64_t per = ((panels + wpg - 1) / wpg + kB - 1) / kB * kB;
  const int64_t qa = g < groups && part * per < panels ? part * per : panels;
  const int64_t qb = qa + per < panels ? qa + per : panels;
  // Lane l loads the ordinal range of panel qq + l once per 64 panels; the
  // panels then go in batches of 8, each batch's sums and row tags loading
  // while the previous batch is added.
  struct Batch {
    uint32_t lo[kB], n[kB];
    double v[kB][2];
    uint16_t r[kB][2];
  };
  const auto load = [&](Batch& b, uint32_t lo_r, uint32_t hi_r, int j0) {
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      b.lo[j] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo_r), j0 + j));
      b.n[j] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hi_r), j0 + j)) -
               b.lo[j];
    }
#pragma unroll
    for (int j = 0; j < kB; ++j) {
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const uint32_t k = static_cast<uint32_t>(lane + u * kWave);
        b.v[j][u] = k < b.n[j] ? partial[b.lo[j] + k] : 0.0;
        b.r[j][u] = k < b.n[j] ? seg_lrow[b.lo[j] + k] : static_cast<uint16_t>(0);
      }
    }
  };
  // Up to two segments per lane per panel come from the batch; longer ranges
  // (more than 128 of the group's rows in one panel) finish here.
  const auto add = [&](const Batch& b) {
#pragma unroll
    for (int j = 0; j < kB; ++j) {
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (static_cast<uint32_t>(lane + u * kWave) < b.n[j])
          __hip_atomic_fetch_add(&acc[b.r[j][u]], b.v[j][u], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
      for (uint32_t k = lane + 2 * kWave; k < b.n[j]; k += kWave)
        __hip_atomic_fetch_add(&acc[seg_lrow[b.lo[j] + k]], partial[b.lo[j] + k], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  };
  for (int64_t qq = qa; qq < qb; qq += kWave) {
    const int64_t q = qq + lane;
    const uint32_t lo_r = q < qb ? ctable[q * stride + g] : 0u;
    const uint32_t hi_r = q < qb ? ctable[q * stride + g + 1] : 0u;
    const int64_t left = qb - qq;
    const int nb = static_cast<int>((left < kWave ? left : kWave) / kB);
    Batch A, B;
    load(A, lo_r, hi_r, 0);
    for (int bi = 0; bi < nb; bi += 2) {
      if (bi + 1 < nb) load(B, lo_r, hi_r, (bi + 1) * kB);
      add(A);
      if (bi + 1 >= nb) break;
      if (bi + 2 < nb) load(A, lo_r, hi_r, (bi + 2) * kB);
      add(B);
    }
  }
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * gpb * kStreamGroupRows;
  for (int k = threadIdx.x; k < gpb * kStreamGroupRows && base + k < nlong; k += kLdsBlock) {
    const int gi = k / kStreamGroupRows, r = k % kStreamGroupRows;
    double sum = 0.0;
    for (int w = gi * wpg; w < gi * wpg + wpg; ++w) sum += sacc[w * kStreamGroupRows + r];
    finish_row<V>(p, long_rows[base + k], Acc::from_f64(sum));
  }
}

// Groups per combine work-group: the most (up to 16, one wave each) that
// still give a work-group per CU.
int combine_groups_per_block(int64_t groups) {
  int gpb = kCombineWaves;
  while (gpb > 1 && (groups + gpb - 1) / gpb < kCombineTargetBlocks) gpb /= 2;
  return gpb;
}

// ---- stream-form analysis -------------------------------------------------
// A listed row's entries in one panel are a contiguous run of its sorted
// columns: a segment.  G lanes walk each listed row, flag the entries that
// start a segment (the row's first entry, or a panel change), and either
// count them (cnt) or record them (start, panel, listed row) at the row's
// base ordinal plus the heads before them.  Segments are in row order here;
// a stable sort by panel then gives the stream order.
constexpr int kSegLanes = 16;

template <typename I, bool FILL>
__global__ __launch_bounds__(kBlock) void stream_segments_kernel(
    const int64_t* __restrict__ offsets, const I* __restrict__ col,
    const uint32_t* __restrict__ long_rows, int64_t nlong, uint32_t panel_cols,
    uint32_t* __restrict__ cnt, const uint32_t* __restrict__ base, uint32_t* __restrict__ sstart,
    uint32_t* __restrict__ spanel, uint32_t* __restrict__ srow) {
  const int64_t i = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / kSegLanes;
  const int sub = threadIdx.x % kSegLanes;
  const int shift = (threadIdx.x % kWave) / kSegLanes * kSegLanes;  // the group's lanes in the wave
  const bool live = i < nlong;  // whole groups agree
  int64_t b = 0, e = 0;
  if (live) {
    const int64_t r = long_rows[i];
    b = offsets[r];
    e = offsets[r + 1];
  }
  uint32_t count = 0;
  const uint32_t at = FILL && live ? base[i] : 0u;
  // Every lane of the wave iterates while any group still has entries.
  int64_t len = e - b;
  int64_t rounds = (len + kSegLanes - 1) / kSegLanes;
  // One shuffle per step, taken by every lane (a shuffle inside a
  // conditional would read inactive lanes as 0).
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const int64_t other = __shfl_xor(rounds, off);
    rounds = rounds > other ? rounds : other;
  }
  for (int64_t t = 0; t < rounds; ++t) {
    const int64_t k = b + t * kSegLanes + sub;
    bool head = false;
    uint32_t q = 0;
    if (k < e) {
      q = static_cast<uint32_t>(static_cast<int64_t>(col[k])) / panel_cols;
      head = k == b || static_cast<uint32_t>(static_cast<int64_t>(col[k - 1])) / panel_cols != q;
    }
    const uint64_t m = (__ballot(head) >> shift) & ((1ull << kSegLanes) - 1ull);
    if (FILL && head) {
      const uint32_t sid = at + count + static_cast<uint32_t>(__popcll(m & ((1ull << sub) - 1ull)));
      sstart[sid] = static_cast<uint32_t>(k);
      spanel[sid] = q;
      srow[sid] = static_cast<uint32_t>(i);
    }
    count += static_cast<uint32_t>(__popcll(m));
  }
  if (!FILL && live && sub == 0) cnt[i] = count;
}

// Segment lengths (row order): up to the next segment of the same row, or
// the row's end.
__global__ __launch_bounds__(kBlock) void stream_seg_len_kernel(
    const int64_t* __restrict__ offsets, const uint32_t* __restrict__ long_rows,
    const uint32_t* __restrict__ sstart, const uint32_t* __restrict__ srow, int64_t nseg,
    uint32_t* __restrict__ slen) {
  const int64_t sid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (sid >= nseg) return;
  const uint32_t i = srow[sid];
  const uint32_t end = sid + 1 < nseg && srow[sid + 1] == i
                           ? sstart[sid + 1]
                           : static_cast<uint32_t>(offsets[long_rows[i] + 1]);
  slen[sid] = end - sstart[sid];
}

// Segment fields in stream order (o: ordinal, sorder[o]: row-order segment).
__global__ __launch_bounds__(kBlock) void stream_sorted_kernel(
    const uint32_t* __restrict__ sorder, const uint32_t* __restrict__ slen,
    const uint32_t* __restrict__ srow, int64_t nseg, uint32_t* __restrict__ lens_s,
    uint32_t* __restrict__ lrow_s) {
  const int64_t o = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (o >= nseg) return;
  const uint32_t sid = sorder[o];
  lens_s[o] = slen[sid];
  lrow_s[o] = srow[sid];
}

// small[q] = first raw stream position of panel q, small[panels] = entries.
__global__ __launch_bounds__(kBlock) void stream_panel_starts_kernel(
    const int64_t* __restrict__ pseg, const uint32_t* __restrict__ dest, int64_t panels,
    int64_t nseg, uint32_t total, uint32_t* __restrict__ small) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q > panels) return;
  const int64_t o = q < panels ? pseg[q] : nseg;
  small[q] = o < nseg ? dest[o] : total;
}

// One thread per segment (stream order): copies its entries into the
// panel's stream (column within the panel, head flag on the first; values
// at their lane-interleaved positions) and records its row within its
// group of kStreamGroupRows listed rows.
template <typename I, typename VB>
__global__ __launch_bounds__(kBlock) void stream_scatter_kernel(
    const I* __restrict__ col, const VB* __restrict__ val, const uint32_t* __restrict__ sorder,
    const uint32_t* __restrict__ sstart, const uint32_t* __restrict__ spanel,
    const uint32_t* __restrict__ lens_s, const uint32_t* __restrict__ lrow_s,
    const uint32_t* __restrict__ dest, const uint32_t* __restrict__ shift, int64_t nseg,
    int64_t panel_cols, uint16_t* __restrict__ scol, VB* __restrict__ sval,
    uint16_t* __restrict__ seg_lrow, uint32_t* __restrict__ seg_src,
    uint32_t* __restrict__ seg_dst, uint32_t* __restrict__ seg_len) {
  const int64_t o = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (o >= nseg) return;
  const uint32_t sid = sorder[o];
  const int64_t src = sstart[sid];
  const int64_t q = spanel[sid];
  const uint32_t len = lens_s[o];
  const int64_t dst = static_cast<int64_t>(dest[o]) + shift[q];
  const int64_t col0 = q * panel_cols;
  seg_lrow[o] = static_cast<uint16_t>(lrow_s[o] % kStreamGroupRows);
  seg_src[o] = static_cast<uint32_t>(src);
  seg_dst[o] = static_cast<uint32_t>(dst);
  seg_len[o] = len;
  for (uint32_t k = 0; k < len; ++k) {
    const uint32_t c = static_cast<uint32_t>(static_cast<int64_t>(col[src + k]) - col0);
    scol[dst + k] = static_cast<uint16_t>(c | (k == 0 ? kStreamHead : 0u));
    sval[stream_value_pos<VB>(dst + k)] = val[src + k];
  }
}

// Values-only refresh of a stream-form plan: one thread per segment copies
// the segment's current values to its lane-interleaved stream positions.
template <typename VB>
__global__ __launch_bounds__(kBlock) void stream_refresh_kernel(
    const VB* __restrict__ val, const uint32_t* __restrict__ seg_src,
    const uint32_t* __restrict__ seg_dst, const uint32_t* __restrict__ seg_len, int64_t nseg,
    VB* __restrict__ sval) {
  const int64_t o = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (o >= nseg) return;
  const int64_t src = seg_src[o];
  const int64_t dst = seg_dst[o];
  const uint32_t len = seg_len[o];
  for (uint32_t k = 0; k < len; ++k) sval[stream_value_pos<VB>(dst + k)] = val[src + k];
}

// Completes slice_info from the host's staging table {start, end, raw start,
// panel | kStreamClosed on the panel's last slice}: the segment open before
// the slice's first entry, and whether the slice's last segment ends in it.
__global__ __launch_bounds__(kBlock) void stream_slice_kernel(
    const uint32_t* __restrict__ stage, int64_t slices, const int64_t* __restrict__ pseg,
    const uint32_t* __restrict__ dest, const uint16_t* __restrict__ scol,
    uint32_t* __restrict__ slice_info) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (k >= slices) return;
  const uint32_t start = stage[4 * k], end = stage[4 * k + 1], raw = stage[4 * k + 2];
  const uint32_t w = stage[4 * k + 3];
  const int64_t q = w & ~kStreamClosed;
  // Last ordinal o of panel q with dest[o] <= raw: the segment holding raw.
  int64_t lo = pseg[q], hi = pseg[q + 1] - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (dest[mid] <= raw) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t ord = static_cast<uint32_t>(lo) - (dest[lo] == raw ? 1u : 0u);
  const bool closed = (w & kStreamClosed) != 0 || (scol[end] & kStreamHead) != 0;
  slice_info[4 * k] = start;
  slice_info[4 * k + 1] = end;
  slice_info[4 * k + 2] = ord;
  slice_info[4 * k + 3] = static_cast<uint32_t>(q) | (closed ? kStreamClosed : 0u);
}

// Combine table: ctable[q*(groups+1) + g] = first ordinal of panel q whose
// listed row is at least g*kStreamGroupRows.
__global__ __launch_bounds__(kBlock) void stream_ctable_kernel(
    const int64_t* __restrict__ pseg, const uint32_t* __restrict__ lrow_s, int64_t panels,
    int64_t groups, uint32_t* __restrict__ ctable) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t stride = groups + 1;
  if (idx >= panels * stride) return;
  const int64_t q = idx / stride, g = idx - q * stride;
  const uint32_t bound = static_cast<uint32_t>(g * kStreamGroupRows);
  int64_t lo = pseg[q], hi = pseg[q + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (lrow_s[mid] < bound) lo = mid + 1;
    else hi = mid;
  }
  ctable[idx] = static_cast<uint32_t>(lo);
}

// ---------------------------------------------------------------------------
// ordered: one lane per row, exact serial arithmetic (FP32 / FP64 only).
// ---------------------------------------------------------------------------
template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmv_ordered_kernel(KParams p) {
#pragma clang fp contract(off)
  const int64_t row =
      xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (row >= p.rows) return;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ x = static_cast<const T*>(p.x);
  T* y = static_cast<T*>(p.y) + row * p.incy;
  const int64_t b = p.offsets[row];
  const int64_t e = p.offsets[row + 1];
  const T alpha = static_cast<T>(p.alpha);
  if (p.form == 0) {
    // reference_blas.cc:22-38
    T product = T(0);
    for (int64_t k = b; k < e; ++k) {
      const T term = val[k] * x[static_cast<int64_t>(col[k]) * p.incx];
      product = product + term;
    }
    T result = alpha * product;
    const T beta = static_cast<T>(p.beta);
    if (beta != T(0)) {
      const T scaled = beta * *y;
      result = result + scaled;
    }
    *y = result;
  } else {
    // blas.h:1653-1665: y += (alpha*v) * x per stored entry, in order.
    T acc = *y;
    for (int64_t k = b; k < e; ++k) {
      const T av = alpha * val[k];
      const T term = av * x[static_cast<int64_t>(col[k]) * p.incx];
      acc = acc + term;
    }
    *y = acc;
  }
}

// Ordered, 16 lanes a row (rows averaging kOrderedGroupMin entries or
// more): the lanes load 16 entries of the row and form their terms at once,
// then every lane adds the 16 terms to the running sum one after another
// (each broadcast across the group by DPP), so the sum keeps the serial
// order of the one-lane kernel above, bit for bit, while the row's loads
// and gathers go out 16 at a time.  The next step's entries load before the
// current step's additions.
constexpr int kOrderedLanes = 16;
constexpr int64_t kOrderedGroupMin = 8;

template <int U, typename T>
__device__ __forceinline__ T ordered_bcast(T v) {
  return __builtin_amdgcn_update_dpp(T(0), v, 0x150 + U, 0xF, 0xF, false);
}

template <typename T, int U>
__device__ __forceinline__ void ordered_add_step(T term, int64_t n, T* acc) {
#pragma clang fp contract(off)
  if constexpr (U < kOrderedLanes) {
    const T t = ordered_bcast<U>(term);
    if (U < n) *acc = *acc + t;
    ordered_add_step<T, U + 1>(term, n, acc);
  }
}

template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmv_ordered16_kernel(KParams p) {
#pragma clang fp contract(off)
  constexpr int G = kOrderedLanes;
  const int64_t row = (xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) / G;
  if (row >= p.rows) return;  // whole group
  const int j = threadIdx.x % G;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ x = static_cast<const T*>(p.x);
  T* y = static_cast<T*>(p.y) + row * p.incy;
  const int64_t b = p.offsets[row];
  const int64_t e = p.offsets[row + 1];
  const T alpha = static_cast<T>(p.alpha);
  // The term of entry k: v*x (compat, reference_blas.cc:22-38) or
  // (alpha*v)*x (standard, blas.h:1653-1665).
  const auto term_of = [&](int64_t k) {
    if (k >= e) return T(0);
    const T xv = x[static_cast<int64_t>(col[k]) * p.incx];
    const T v = p.form == 0 ? val[k] : alpha * val[k];
    return v * xv;
  };
  T acc = p.form == 0 ? T(0) : *y;
  T term = term_of(b + j);
  for (int64_t t = b; t < e; t += G) {
    const T cur = term;
    term = term_of(t + G + j);
    ordered_add_step<T, 0>(cur, e - t, &acc);
  }
  if (j == 0) {
    if (p.form == 0) {
      T result = alpha * acc;
      const T beta = static_cast<T>(p.beta);
      if (beta != T(0)) {
        const T scaled = beta * *y;
        result = result + scaled;
      }
      *y = result;
    } else {
      *y = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// merge-path, wave-level and persistent
//
// The merge sequence (row ends interleaved with nonzeros, Merrill & Garland)
// is cut into wave tiles of kWTile items.  Each persistent wave owns a
// contiguous run of tiles and walks them in order:
//   - the next tile's column indices, values and row ends are loaded into
//     registers while the current tile is reduced (one-tile prefetch), and
//     no block barrier stands in between, so those loads stay in flight;
//   - products and row ends are staged in the wave's own LDS slice, each
//     lane walks kWIpt consecutive merge items, and a wave64 segmented scan
//     hands partial rows to the lane that finishes them;
//   - the open row at a tile's end carries into the next tile in registers;
//     only the carry left at a wave's end goes to the fix-up kernel.
// ---------------------------------------------------------------------------
constexpr int kWIpt = 8;                         // merge items per lane
constexpr int kWTile = kWave * kWIpt;            // merge items per wave tile
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int64_t kMaxMergeWaves = 65536;        // carry slots in the workspace

// Tile start coordinates on the merge path: coords[2t] = row, [2t+1] = nz.
__global__ __launch_bounds__(kBlock) void merge_search_kernel(
    const int64_t* __restrict__ offsets, int64_t rows, int64_t nnz,
    int64_t ntiles, int64_t* __restrict__ coords) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t > ntiles) return;
  const int64_t total = rows + nnz;
  const int64_t d = t * kWTile < total ? t * kWTile : total;
  int64_t lo = d - nnz > 0 ? d - nnz : 0;
  int64_t hi = d < rows ? d : rows;
  // Largest prefix of row-end items that precede nz item (d - lo - 1).
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (offsets[mid + 1] <= d - mid - 1) lo = mid + 1;
    else hi = mid;
  }
  coords[2 * t] = lo;
  coords[2 * t + 1] = d - lo;
}

template <typename V, typename I>
struct TileRegs {
  I c[kWIpt];
  V v[kWIpt];
  int32_t e[kWIpt];
  int32_t e_last;  // row end for tile row kWTile (only when nrows == kWTile)
};

// Row end of tile row r (0 <= r <= nrows) relative to the tile's first nz;
// the open row's end is clamped (only "beyond the tile" matters).
__device__ __forceinline__ int32_t rel_row_end(const KParams& p, int64_t i0,
                                               int64_t j0, int r) {
  const int64_t gr = i0 + r;
  int64_t rel = (gr < p.rows ? ld_stream(p.offsets + gr + 1) : p.nnz) - j0;
  return static_cast<int32_t>(rel > kWTile + 1 ? kWTile + 1 : rel);
}

template <typename V, typename I>
__device__ __forceinline__ void load_tile(const KParams& p, TileRegs<V, I>& t,
                                          int64_t i0, int64_t j0, int64_t i1,
                                          int64_t j1, int lane) {
  const V* __restrict__ val = static_cast<const V*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const int nnzt = static_cast<int>(j1 - j0);
  const int nrows = static_cast<int>(i1 - i0);
#pragma unroll
  for (int u = 0; u < kWIpt; ++u) {
    const int k = lane + u * kWave;
    if (k < nnzt) {
      t.c[u] = ld_stream(col + j0 + k);
      t.v[u] = ld_stream(val + j0 + k);
    }
  }
#pragma unroll
  for (int u = 0; u < kWIpt; ++u) {
    const int r = lane + u * kWave;
    if (r <= nrows) t.e[u] = rel_row_end(p, i0, j0, r);
  }
  if (lane == 0 && nrows == kWTile) t.e_last = rel_row_end(p, i0, j0, kWTile);
}

template <typename V, typename I, bool TWOSUM>
__global__ __launch_bounds__(kBlock) void spmv_merge_kernel(
    KParams p, const int64_t* __restrict__ coords, int64_t ntiles,
    int64_t tiles_per_wave, int64_t* __restrict__ carry_rows,
    void* __restrict__ carry_vals) {
  using A = typename acc_of<V>::type;
  using Acc = typename AccSel<V, TWOSUM>::type;
  __shared__ int32_t s_end_all[kWavesPerBlock][kWTile + 1];
  __shared__ A s_val_all[kWavesPerBlock][kWTile];
  __shared__ float s_err_all[kWavesPerBlock][TWOSUM ? kWTile : 1];
  const int lane = threadIdx.x % kWave;
  const int wib = threadIdx.x / kWave;
  int32_t* s_end = s_end_all[wib];
  A* s_val = s_val_all[wib];
  float* s_err = s_err_all[wib];
  const int64_t wave = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + wib;
  const int64_t t_begin = wave * tiles_per_wave;
  int64_t t_end = t_begin + tiles_per_wave;
  if (t_end > ntiles) t_end = ntiles;
  if (t_begin >= t_end) {
    if (lane == 0) carry_rows[wave] = -1;
    return;
  }
  const A* __restrict__ x = static_cast<const A*>(p.x);

  int64_t i0 = coords[2 * t_begin], j0 = coords[2 * t_begin + 1];
  int64_t i1 = coords[2 * t_begin + 2], j1 = coords[2 * t_begin + 3];
  TileRegs<V, I> cur, nxt;
  load_tile(p, cur, i0, j0, i1, j1, lane);
  Acc carry = Acc::zero();  // partial of row i0 from earlier tiles of this wave

  for (int64_t t = t_begin; t < t_end; ++t) {
    const int nnzt = static_cast<int>(j1 - j0);
    const int nrows = static_cast<int>(i1 - i0);
    const int items = nrows + nnzt;
    // Gathers for this tile.
    A xv[kWIpt];
#pragma unroll
    for (int u = 0; u < kWIpt; ++u) {
      const int k = lane + u * kWave;
      if (k < nnzt) xv[u] = x[static_cast<int64_t>(cur.c[u]) * p.incx];
    }
    // Prefetch the next tile while this one is reduced.
    int64_t ni1 = i1, nj1 = j1;
    if (t + 1 < t_end) {
      ni1 = coords[2 * t + 4];
      nj1 = coords[2 * t + 5];
      load_tile(p, nxt, i1, j1, ni1, nj1, lane);
    }
    // Stage products and row ends in this wave's LDS slice.
#pragma unroll
    for (int u = 0; u < kWIpt; ++u) {
      const int k = lane + u * kWave;
      if (k < nnzt) {
        const A v = widen<A>(cur.v[u]);
        if constexpr (TWOSUM) {
#pragma clang fp contract(off)
          const float prod = v * xv[u];
          s_val[k] = prod;
          s_err[k] = __builtin_fmaf(v, xv[u], -prod);
        } else {
          s_val[k] = v * xv[u];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kWIpt; ++u) {
      const int r = lane + u * kWave;
      if (r <= nrows) s_end[r] = cur.e[u];
    }
    if (lane == 0 && nrows == kWTile) s_end[kWTile] = cur.e_last;
    __builtin_amdgcn_wave_barrier();
    // Lane-level merge walk of kWIpt items.
    const int d = lane * kWIpt < items ? lane * kWIpt : items;
    int lo = d - nnzt > 0 ? d - nnzt : 0;
    int hi = d < nrows ? d : nrows;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_end[mid] <= d - mid - 1) lo = mid + 1;
      else hi = mid;
    }
    int r = lo, j = d - lo;
    const int dend = d + kWIpt < items ? d + kWIpt : items;
    Acc acc = Acc::zero();
    int first_done = -1;
    for (int step = d; step < dend; ++step) {
      if (j < s_end[r]) {
        if constexpr (TWOSUM) acc.add_pair(s_val[j], s_err[j]);
        else acc.add(s_val[j]);
        ++j;
      } else {
        s_val[nnzt + r] = acc.s;
        if constexpr (TWOSUM) s_err[nnzt + r] = acc.c;
        if (first_done < 0) first_done = r;
        acc = Acc::zero();
        ++r;
      }
    }
    // Segmented inclusive scan over lanes; keys (open rows) never decrease.
    const int key = r;
    Acc run = acc;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int k_up = __shfl_up(key, off);
      const Acc a_up = run.shfl_up(off);
      if (lane >= off && k_up == key) {
        Acc tt = a_up;
        tt.merge_after(run);
        run = tt;
      }
    }
    // Prefix owed to this lane's first finished row: the predecessor's run,
    // preceded by the carry from earlier tiles when that run is for row 0.
    int prev_key = __shfl_up(key, 1);
    Acc prev = run.shfl_up(1);
    if (lane == 0) {
      prev_key = 0;
      prev = carry;
    } else if (prev_key == 0) {
      Acc tt = carry;
      tt.merge_after(prev);
      prev = tt;
    }
    if (first_done >= 0 && prev_key == first_done) {
      Acc mine = Acc::zero();
      mine.s = s_val[nnzt + first_done];
      if constexpr (TWOSUM) mine.c = s_err[nnzt + first_done];
      Acc tt = prev;
      tt.merge_after(mine);
      s_val[nnzt + first_done] = tt.s;
      if constexpr (TWOSUM) s_err[nnzt + first_done] = tt.c;
    }
    // Open row at the tile end: the last lane's run (plus the old carry when
    // no row finished in this tile).
    Acc last;
    last.s = __shfl(run.s, kWave - 1);
    if constexpr (TWOSUM) last.c = __shfl(run.c, kWave - 1);
    if (nrows == 0) {
      Acc tt = carry;
      tt.merge_after(last);
      carry = tt;
    } else {
      carry = last;
    }
    __builtin_amdgcn_wave_barrier();
    // Coalesced epilogue for rows finished in this tile.
#pragma unroll 1
    for (int rr = lane; rr < nrows; rr += kWave) {
      Acc a = Acc::zero();
      a.s = s_val[nnzt + rr];
      if constexpr (TWOSUM) a.c = s_err[nnzt + rr];
      finish_row<V>(p, i0 + rr, a);
    }
    __builtin_amdgcn_wave_barrier();
    // Advance.
    i0 = i1;
    j0 = j1;
    i1 = ni1;
    j1 = nj1;
    cur = nxt;
  }
  // Carry left at this wave's end: the open row, if this wave read any of it.
  if (lane == 0) {
    const int64_t open_begin = i0 < p.rows ? p.offsets[i0] : p.nnz;
    if (i0 < p.rows && j0 > open_begin) {
      carry_rows[wave] = i0;
      static_cast<Acc*>(carry_vals)[wave] = carry;
    } else {
      carry_rows[wave] = -1;
    }
  }
}

// Adds wave carries to rows that crossed wave boundaries.  The first wave of
// each run of equal carry rows sums the run in order and adds it once.
template <typename V, bool TWOSUM>
__global__ __launch_bounds__(kBlock) void spmv_merge_fixup_kernel(
    KParams p, int64_t nwaves, const int64_t* __restrict__ carry_rows,
    const void* __restrict__ carry_vals) {
  using Acc = typename AccSel<V, TWOSUM>::type;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= nwaves) return;
  const int64_t row = carry_rows[t];
  if (row < 0) return;
  if (t > 0 && carry_rows[t - 1] == row) return;
  const Acc* vals = static_cast<const Acc*>(carry_vals);
  Acc sum = vals[t];
  for (int64_t u = t + 1; u < nwaves && carry_rows[u] == row; ++u)
    sum.merge_after(vals[u]);
  add_carry<V>(p, row, sum);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
int64_t merge_tiles(const SpmvArgs& a) {
  const int64_t total = a.a.outer + a.a.nnz;
  return (total + kWTile - 1) / kWTile;
}

int device_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

// Persistent grid for the merge kernel: as many resident blocks as the
// kernel's occupancy allows, so every wave runs exactly one tile range.
template <typename V, typename I, bool TWOSUM>
int64_t merge_blocks() {
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, spmv_merge_kernel<V, I, TWOSUM>, kBlock, 0) != hipSuccess || nb <= 0)
      nb = 4;
    per_cu = nb;
  }
  int64_t blocks = static_cast<int64_t>(per_cu) * device_cus();
  if (blocks * kWavesPerBlock > kMaxMergeWaves) blocks = kMaxMergeWaves / kWavesPerBlock;
  return blocks;
}

template <typename V, bool TWOSUM>
size_t carry_bytes() {
  return sizeof(typename AccSel<V, TWOSUM>::type);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Short rows of a skewed matrix are bound by x gathers from the memory-side
// cache, the stream-form panel kernels by HBM streaming: the short rows run
// on the side stream while the panel kernels run on the main stream.
template <typename V, typename I>
hipError_t launch_short_rows(const SpmvArgs& a, const KParams& kp, int64_t skip, hipStream_t s) {
  const int64_t rows = a.a.outer;
  switch (a.panel.short_alg) {
#define ASC_SIDE_CASE(ALG, G)                                                           \
  case SpmvAlg::ALG:                                                                    \
    spmv_vector_kernel<G, V, I, false>                                                  \
        <<<static_cast<unsigned>((rows + kBlock / G - 1) / (kBlock / G)), kBlock, 0, s>>>( \
            kp, skip);                                                                  \
    return hipGetLastError();
    ASC_SIDE_CASE(Vec1, 1)
    ASC_SIDE_CASE(Vec2, 2)
    ASC_SIDE_CASE(Vec4, 4)
    ASC_SIDE_CASE(Vec8, 8)
    ASC_SIDE_CASE(Vec16, 16)
    ASC_SIDE_CASE(Vec32, 32)
    ASC_SIDE_CASE(Vec64, 64)
#undef ASC_SIDE_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <typename V, typename I>
hipError_t launch_stream_forked(const SpmvArgs& a, const KParams& kp, void* ws, hipStream_t s,
                                const SideStream& side) {
  const PanelPlan& pp = a.panel;
  double* carry = static_cast<double*>(ws);
  // Rows below long_min (at the default of 1: only empty rows) are finished
  // by the short-row pass on the side stream; with every row listed there is
  // nothing to fork.
  const bool shorts = pp.nlong < a.a.outer;
  hipError_t e = shorts ? hipEventRecord(side.fork, s) : hipSuccess;
  if (e != hipSuccess) return e;
  spmv_stream_kernel<V><<<static_cast<unsigned>(pp.wgs), kStreamBlock, 0, s>>>(
      static_cast<const typename acc_of<V>::type*>(kp.x), kp.incx, a.a.inner, pp.panel_cols,
      pp.scol, static_cast<const V*>(pp.sval), pp.slice_info, pp.wg_table, pp.spartial, carry,
      a.spin_limit, a.fault);
  e = hipGetLastError();
  // The side stream's work is joined below whatever happens after the fork.
  hipError_t es = hipSuccess, ej = hipSuccess;
  if (shorts) {
    es = hipStreamWaitEvent(side.stream, side.fork, 0);
    if (es == hipSuccess) es = launch_short_rows<V, I>(a, kp, pp.long_min, side.stream);
    ej = hipEventRecord(side.join, side.stream);
  }
  if (e == hipSuccess) {
    spmv_stream_fixup_kernel<<<static_cast<unsigned>((pp.slices + kBlock - 1) / kBlock), kBlock, 0,
                               s>>>(pp.slice_info, pp.slices, carry, pp.spartial);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const int gpb = combine_groups_per_block(pp.groups);
    spmv_stream_combine_kernel<V>
        <<<static_cast<unsigned>((pp.groups + gpb - 1) / gpb), kLdsBlock, 0, s>>>(
            kp, pp.long_rows, pp.nlong, pp.panels, pp.ctable, pp.groups, pp.seg_lrow, pp.spartial,
            gpb);
    e = hipGetLastError();
  }
  const hipError_t ew =
      !shorts ? hipSuccess : ej == hipSuccess ? hipStreamWaitEvent(s, side.join, 0) : ej;
  if (e != hipSuccess) return e;
  if (es != hipSuccess) return es;
  return ew;
}

template <typename V, typename I, bool TWOSUM>
hipError_t launch_typed(const SpmvArgs& a, const KParams& kp, void* ws,
                        hipStream_t s, const SideStream* side) {
  const int64_t rows = a.a.outer;
  auto blocks_for = [](int64_t n, int64_t per) {
    return static_cast<unsigned>((n + per - 1) / per);
  };
  switch (a.alg) {
    case SpmvAlg::Ordered: {
      if constexpr (!TWOSUM && (std::is_same<V, double>::value ||
                                std::is_same<V, float>::value)) {
        if (a.a.nnz >= kOrderedGroupMin * rows)
          spmv_ordered16_kernel<V, I><<<blocks_for(rows, kBlock / kOrderedLanes), kBlock, 0, s>>>(kp);
        else
          spmv_ordered_kernel<V, I><<<blocks_for(rows, kBlock), kBlock, 0, s>>>(kp);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    }
    case SpmvAlg::Merge: {
      const int64_t nt = merge_tiles(a);
      char* w = static_cast<char*>(ws);
      int64_t* coords = reinterpret_cast<int64_t*>(w);
      w += align256(sizeof(int64_t) * 2 * (nt + 1));
      int64_t* crow = reinterpret_cast<int64_t*>(w);
      w += align256(sizeof(int64_t) * kMaxMergeWaves);
      void* cval = w;
      // Waves: one per resident slot, each with a contiguous run of tiles.
      int64_t blocks = merge_blocks<V, I, TWOSUM>();
      int64_t waves = blocks * kWavesPerBlock;
      if (waves > nt) waves = nt;
      const int64_t per_wave = (nt + waves - 1) / waves;
      waves = (nt + per_wave - 1) / per_wave;
      blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
      const int64_t slots = blocks * kWavesPerBlock;
      merge_search_kernel<<<blocks_for(nt + 1, kBlock), kBlock, 0, s>>>(
          a.a.offsets, rows, a.a.nnz, nt, coords);
      spmv_merge_kernel<V, I, TWOSUM><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
          kp, coords, nt, per_wave, crow, cval);
      spmv_merge_fixup_kernel<V, TWOSUM><<<blocks_for(slots, kBlock), kBlock, 0, s>>>(
          kp, slots, crow, cval);
      return hipGetLastError();
    }
#define ASC_VEC_CASE(ALG, G)                                                  \
  case SpmvAlg::ALG:                                                          \
    spmv_vector_kernel<G, V, I, TWOSUM>                                       \
        <<<blocks_for(rows, kBlock / G), kBlock, 0, s>>>(kp, INT64_MAX);      \
    return hipGetLastError();
    ASC_VEC_CASE(Vec1, 1)
    ASC_VEC_CASE(Vec2, 2)
    ASC_VEC_CASE(Vec4, 4)
    ASC_VEC_CASE(Vec8, 8)
    ASC_VEC_CASE(Vec16, 16)
    ASC_VEC_CASE(Vec32, 32)
    ASC_VEC_CASE(Vec64, 64)
#undef ASC_VEC_CASE
    case SpmvAlg::Panel: {
      const PanelPlan& pp = a.panel;
      const int64_t skip = pp.nlong > 0 ? pp.long_min : INT64_MAX;
      hipError_t e = hipErrorInvalidValue;
      if constexpr (!TWOSUM) {
        if (pp.stream && pp.nlong > 0 && side != nullptr && side->stream != nullptr)
          return launch_stream_forked<V, I>(a, kp, ws, s, *side);
      }
      switch (pp.short_alg) {
#define ASC_SHORT_CASE(ALG, G)                                                \
  case SpmvAlg::ALG:                                                          \
    spmv_vector_kernel<G, V, I, TWOSUM>                                       \
        <<<blocks_for(rows, kBlock / G), kBlock, 0, s>>>(kp, skip);           \
    e = hipGetLastError();                                                    \
    break;
        ASC_SHORT_CASE(Vec1, 1)
        ASC_SHORT_CASE(Vec2, 2)
        ASC_SHORT_CASE(Vec4, 4)
        ASC_SHORT_CASE(Vec8, 8)
        ASC_SHORT_CASE(Vec16, 16)
        ASC_SHORT_CASE(Vec32, 32)
        ASC_SHORT_CASE(Vec64, 64)
#undef ASC_SHORT_CASE
        default:
          break;
      }
      if (e != hipSuccess || pp.nlong == 0) return e;
      if (pp.stream) {
        if constexpr (TWOSUM) {
          return hipErrorInvalidValue;  // the stream form accumulates plainly
        } else {
          double* carry = static_cast<double*>(ws);
          spmv_stream_kernel<V><<<static_cast<unsigned>(pp.wgs), kStreamBlock, 0, s>>>(
              static_cast<const typename acc_of<V>::type*>(kp.x), kp.incx, a.a.inner,
              pp.panel_cols, pp.scol, static_cast<const V*>(pp.sval), pp.slice_info,
              pp.wg_table, pp.spartial, carry, a.spin_limit, a.fault);
          e = hipGetLastError();
          if (e != hipSuccess) return e;
          spmv_stream_fixup_kernel<<<blocks_for(pp.slices, kBlock), kBlock, 0, s>>>(
              pp.slice_info, pp.slices, carry, pp.spartial);
          e = hipGetLastError();
          if (e != hipSuccess) return e;
          const int gpb = combine_groups_per_block(pp.groups);
          spmv_stream_combine_kernel<V>
              <<<static_cast<unsigned>((pp.groups + gpb - 1) / gpb), kLdsBlock, 0, s>>>(
                  kp, pp.long_rows, pp.nlong, pp.panels, pp.ctable, pp.groups, pp.seg_lrow,
                  pp.spartial, gpb);
          return hipGetLastError();
        }
      }
      double* partial = static_cast<double*>(ws);
      const int64_t groups = (pp.panels + 7) / 8;
      const int64_t chunk_rows = (pp.nlong + pp.chunks - 1) / pp.chunks;
      switch (pp.long_alg) {
#define ASC_LONG_CASE(ALG, G)                                                 \
  case SpmvAlg::ALG: {                                                        \
    if (pp.lds) {                                                             \
      spmv_panel_lds_kernel<G, V, I, TWOSUM>                                  \
          <<<static_cast<unsigned>(pp.panels * pp.chunks), kLdsBlock, 0, s>>>( \
              kp, a.a.inner, pp.row_begin, pp.splits, pp.nlong, pp.panels,    \
              pp.panel_cols, chunk_rows, partial);                            \
      e = hipGetLastError();                                                  \
      break;                                                                  \
    }                                                                         \
    const int64_t row_blocks = (pp.nlong + kBlock / G - 1) / (kBlock / G);    \
    const int64_t rb_grid = row_blocks < kPanelRowBlockCap / groups           \
                                ? row_blocks                                  \
                                : kPanelRowBlockCap / groups;                 \
    spmv_panel_kernel<G, V, I, TWOSUM>                                        \
        <<<static_cast<unsigned>(8 * groups * rb_grid), kBlock, 0, s>>>(      \
            kp, pp.row_begin, pp.splits, pp.nlong, pp.panels, row_blocks,     \
            rb_grid, partial);                                                \
    e = hipGetLastError();                                                    \
    break;                                                                    \
  }
        ASC_LONG_CASE(Vec1, 1)
        ASC_LONG_CASE(Vec2, 2)
        ASC_LONG_CASE(Vec4, 4)
        ASC_LONG_CASE(Vec8, 8)
        ASC_LONG_CASE(Vec16, 16)
        ASC_LONG_CASE(Vec32, 32)
        ASC_LONG_CASE(Vec64, 64)
#undef ASC_LONG_CASE
        default:
          e = hipErrorInvalidValue;
          break;
      }
      if (e != hipSuccess) return e;
      spmv_panel_combine_kernel<V, TWOSUM><<<blocks_for(pp.nlong, kBlock), kBlock, 0, s>>>(
          kp, pp.long_rows, pp.nlong, pp.panels, partial);
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}

template <typename V, bool TWOSUM>
hipError_t launch_by_index(const SpmvArgs& a, const KParams& kp, void* ws,
                           hipStream_t s, const SideStream* side) {
  if (a.a.index_bytes == 4) return launch_typed<V, int32_t, TWOSUM>(a, kp, ws, s, side);
  return launch_typed<V, int64_t, TWOSUM>(a, kp, ws, s, side);
}

template <typename V>
size_t ws_typed(const SpmvArgs& a) {
  if (a.alg == SpmvAlg::Panel && a.panel.stream)  // one carry per slice
    return align256(sizeof(double) * static_cast<size_t>(a.panel.slices > 0 ? a.panel.slices : 1));
  if (a.alg == SpmvAlg::Panel)  // FP64 partial sum per (panel, long row)
    return align256(sizeof(double) * static_cast<size_t>(a.panel.panels) *
                    static_cast<size_t>(a.panel.nlong));
  if (a.alg != SpmvAlg::Merge) return 0;
  const int64_t nt = merge_tiles(a);
  const size_t cb = a.twosum ? carry_bytes<V, true>() : carry_bytes<V, false>();
  return align256(sizeof(int64_t) * 2 * (nt + 1)) +
         align256(sizeof(int64_t) * kMaxMergeWaves) + align256(cb * kMaxMergeWaves);
}

}  // namespace

size_t elem_bytes(Elem e) {
  switch (e) {
    case Elem::F64: return 8;
    case Elem::F32: return 4;
    default: return 2;
  }
}

size_t spmv_workspace_bytes(const SpmvArgs& a) {
  switch (a.a.elem) {
    case Elem::F64: return ws_typed<double>(a);
    case Elem::F32: return ws_typed<float>(a);
    case Elem::BF16: return ws_typed<bf16_t>(a);
    case Elem::F16: return ws_typed<f16_t>(a);
  }
  return 0;
}

size_t panel_analysis_scratch_bytes(int64_t outer) {
  const size_t n = static_cast<size_t>(outer > 0 ? outer : 1);
  return 2 * align256(sizeof(uint32_t) * n) + align256(sizeof(uint32_t)) +
         align256(scan_u32_scratch_bytes(outer));
}

hipError_t panel_analysis_launch(const Compressed& a, int64_t panel_cols, int64_t panels,
                                 int64_t long_min, uint32_t* long_rows, int64_t* row_begin,
                                 uint32_t* splits, void* scratch, int64_t* out_nlong,
                                 hipStream_t s) {
  *out_nlong = 0;
  const int64_t rows = a.outer;
  if (rows == 0) return hipSuccess;
  char* w = static_cast<char*>(scratch);
  uint32_t* flags = reinterpret_cast<uint32_t*>(w);
  w += align256(sizeof(uint32_t) * rows);
  uint32_t* pos = reinterpret_cast<uint32_t*>(w);
  w += align256(sizeof(uint32_t) * rows);
  uint32_t* count = reinterpret_cast<uint32_t*>(w);
  w += align256(sizeof(uint32_t));
  const unsigned blocks = static_cast<unsigned>((rows + kBlock - 1) / kBlock);
  panel_flag_kernel<<<blocks, kBlock, 0, s>>>(a.offsets, rows, long_min, flags);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = scan_u32_exclusive(flags, pos, rows, w, s);
  if (e == hipSuccess) {
    panel_list_kernel<<<blocks, kBlock, 0, s>>>(flags, pos, rows, long_rows, count);
    e = hipGetLastError();
  }
  uint32_t host_count = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&host_count, count, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const int64_t nlong = host_count;
  *out_nlong = nlong;
  if (nlong == 0) return hipSuccess;
  const unsigned lblocks = static_cast<unsigned>((nlong + kBlock - 1) / kBlock);
  if (a.index_bytes == 4)
    panel_split_kernel<int32_t><<<lblocks, kBlock, 0, s>>>(
        a.offsets, static_cast<const int32_t*>(a.indices), long_rows, nlong, panel_cols,
        panels, row_begin, splits);
  else
    panel_split_kernel<int64_t><<<lblocks, kBlock, 0, s>>>(
        a.offsets, static_cast<const int64_t*>(a.indices), long_rows, nlong, panel_cols,
        panels, row_begin, splits);
  return hipGetLastError();
}

StreamLayout stream_layout(int64_t rows, int64_t nlong, int64_t nnz_long, int64_t segs_cap,
                           int64_t panels, size_t value_bytes) {
  StreamLayout L;
  L.nlong = nlong;
  L.nnz_long = nnz_long;
  L.panels = panels;
  L.groups = (nlong + kStreamGroupRows - 1) / kStreamGroupRows;
  L.entries_cap = nnz_long + panels * kStreamTile;  // entries plus per-panel padding
  L.slices_cap = L.entries_cap / kStreamSlice + panels + 1;
  L.segs_cap = segs_cap;
  const size_t segs = static_cast<size_t>(segs_cap > 0 ? segs_cap : 1);
  const size_t n = static_cast<size_t>(nlong > 0 ? nlong : 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align256(bytes > 0 ? bytes : 1);
    return at;
  };
  L.long_rows = take(sizeof(uint32_t) * static_cast<size_t>(rows > 0 ? rows : 1));
  L.partial = take(sizeof(double) * segs);
  L.scol = take(sizeof(uint16_t) * static_cast<size_t>(L.entries_cap));
  L.sval = take(value_bytes * static_cast<size_t>(L.entries_cap));
  L.seg_lrow = take(sizeof(uint16_t) * segs);
  L.ctable = take(sizeof(uint32_t) * static_cast<size_t>(panels * (L.groups + 1)));
  L.slice_info = take(4 * sizeof(uint32_t) * static_cast<size_t>(L.slices_cap));
  L.wg_table = take(2 * sizeof(uint32_t) * static_cast<size_t>(L.slices_cap));
  L.seg_src = take(sizeof(uint32_t) * segs);
  L.seg_dst = take(sizeof(uint32_t) * segs);
  L.seg_len = take(sizeof(uint32_t) * segs);
  L.kept_bytes = off;
  L.row_begin = take(sizeof(int64_t) * n);
  L.cnt = take(sizeof(uint32_t) * n);
  L.sbase = take(sizeof(uint32_t) * n);
  L.sstart = take(sizeof(uint32_t) * segs);
  L.spanel = take(sizeof(uint32_t) * segs);
  L.srow = take(sizeof(uint32_t) * segs);
  L.slen = take(sizeof(uint32_t) * segs);
  L.sorder = take(sizeof(uint32_t) * segs);
  L.lens_s = take(sizeof(uint32_t) * segs);
  L.lrow_s = take(sizeof(uint32_t) * segs);
  L.dest = take(sizeof(uint32_t) * segs);
  L.pseg = take(sizeof(int64_t) * static_cast<size_t>(panels + 1));
  L.small = take(sizeof(uint32_t) * static_cast<size_t>(2 * panels + 4));
  L.slice_stage = take(4 * sizeof(uint32_t) * static_cast<size_t>(L.slices_cap));
  const int64_t scan_n = segs_cap > nlong ? segs_cap : nlong;
  L.scan = take(scan_u32_scratch_bytes(scan_n));
  L.sort = take(sort_by_key_workspace_bytes(segs_cap, panels));
  L.list_scratch = take(panel_analysis_scratch_bytes(rows));
  L.total = off;
  return L;
}

hipError_t panel_stream_analysis_launch(const Compressed& a, const StreamLayout& L, void* base,
                                        PanelPlan* pp, hipStream_t s) {
  char* w = static_cast<char*>(base);
  auto at = [&](size_t o) { return reinterpret_cast<uint32_t*>(w + o); };
  uint32_t* long_rows = at(L.long_rows);
  double* partial = reinterpret_cast<double*>(w + L.partial);
  uint16_t* scol = reinterpret_cast<uint16_t*>(w + L.scol);
  void* sval = w + L.sval;
  uint16_t* seg_lrow = reinterpret_cast<uint16_t*>(w + L.seg_lrow);
  uint32_t* ctable = at(L.ctable);
  uint32_t* slice_info = at(L.slice_info);
  uint32_t* wg_table = at(L.wg_table);
  uint32_t *cnt = at(L.cnt), *sbase = at(L.sbase), *sstart = at(L.sstart), *spanel = at(L.spanel),
           *srow = at(L.srow), *slen = at(L.slen), *sorder = at(L.sorder), *lens_s = at(L.lens_s),
           *lrow_s = at(L.lrow_s), *dest = at(L.dest), *small = at(L.small),
           *stage = at(L.slice_stage);
  int64_t* pseg = reinterpret_cast<int64_t*>(w + L.pseg);
  const int64_t panels = pp->panels;
  const uint32_t pcols = static_cast<uint32_t>(pp->panel_cols);
  const size_t vb = elem_bytes(a.elem);
  pp->stream = true;
  pp->long_rows = long_rows;
  pp->row_begin = nullptr;
  pp->splits = nullptr;
  pp->scol = scol;
  pp->sval = sval;
  pp->seg_lrow = seg_lrow;
  pp->ctable = ctable;
  pp->groups = L.groups;
  pp->slice_info = slice_info;
  pp->wg_table = wg_table;
  pp->spartial = partial;
  pp->values_src = a.values;
  pp->stream_base = base;
  pp->stream_bytes = L.kept_bytes;
  pp->seg_src = at(L.seg_src);
  pp->seg_dst = at(L.seg_dst);
  pp->seg_len = at(L.seg_len);
  pp->nseg = 0;
  pp->slices = 0;
  pp->wgs = 0;
  // Listed rows (the row-begin and split outputs of the listing are unused
  // here: one panel of the whole width keeps the split table to one entry
  // per row).
  int64_t nlong = 0;
  hipError_t e = panel_analysis_launch(a, a.inner > 0 ? a.inner : 1, 1, pp->long_min, long_rows,
                                       reinterpret_cast<int64_t*>(w + L.row_begin), cnt,
                                       w + L.list_scratch, &nlong, s);
  if (e != hipSuccess) return e;
  if (nlong != L.nlong) return hipErrorInvalidValue;  // the structure changed since sizing
  pp->nlong = nlong;
  if (nlong == 0) return hipSuccess;
  // Segments in row order: count, scan, record, measure.
  const unsigned lb = static_cast<unsigned>((nlong * kSegLanes + kBlock - 1) / kBlock);
#define ASC_SEGS(I, FILL)                                                                      \
  stream_segments_kernel<I, FILL><<<lb, kBlock, 0, s>>>(                                     \
      a.offsets, static_cast<const I*>(a.indices), long_rows, nlong, pcols, cnt, sbase, sstart, \
      spanel, srow)
  if (a.index_bytes == 4) ASC_SEGS(int32_t, false);
  else ASC_SEGS(int64_t, false);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32_exclusive(cnt, sbase, nlong, w + L.scan, s)) != hipSuccess) return e;
  uint32_t hcount[2] = {0, 0};
  e = hipMemcpyAsync(&hcount[0], sbase + nlong - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&hcount[1], cnt + nlong - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const int64_t nseg = static_cast<int64_t>(hcount[0]) + hcount[1];
  if (nseg > L.segs_cap) return hipErrorInvalidValue;
  if (a.index_bytes == 4) ASC_SEGS(int32_t, true);
  else ASC_SEGS(int64_t, true);
#undef ASC_SEGS
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const unsigned sb = static_cast<unsigned>((nseg + kBlock - 1) / kBlock);
  stream_seg_len_kernel<<<sb, kBlock, 0, s>>>(a.offsets, long_rows, sstart, srow, nseg, slen);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // Stream order: a stable sort of the segments by panel.
  if ((e = sort_by_key_launch(spanel, nseg, panels, sorder, pseg, w + L.sort, s)) != hipSuccess)
    return e;
  stream_sorted_kernel<<<sb, kBlock, 0, s>>>(sorder, slen, srow, nseg, lens_s, lrow_s);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32_exclusive(lens_s, dest, nseg, w + L.scan, s)) != hipSuccess) return e;
  stream_panel_starts_kernel<<<static_cast<unsigned>((panels + 1 + kBlock - 1) / kBlock), kBlock, 0,
                               s>>>(pseg, dest, panels, nseg, static_cast<uint32_t>(L.nnz_long),
                                    small);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  std::vector<uint32_t> hs(static_cast<size_t>(panels + 1));
  e = hipMemcpyAsync(hs.data(), small, sizeof(uint32_t) * hs.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  // Padded panel starts, slices and work-groups (host: panels are few).
  std::vector<uint32_t> shift(static_cast<size_t>(panels));
  std::vector<uint32_t> hstage;
  std::vector<std::vector<uint32_t>> xcd_wgs(kXcds);
  int64_t padded = 0, slices = 0;
  for (int64_t q = 0; q < panels; ++q) {
    const int64_t raw = hs[q];
    const int64_t len = static_cast<int64_t>(hs[q + 1]) - raw;
    shift[q] = static_cast<uint32_t>(padded - raw);
    const int64_t ns = (len + kStreamSlice - 1) / kStreamSlice;
    for (int64_t j = 0; j < ns; ++j) {
      const int64_t b = j * kStreamSlice;
      const int64_t e2 = b + kStreamSlice < len ? b + kStreamSlice : len;
      hstage.push_back(static_cast<uint32_t>(padded + b));
      hstage.push_back(static_cast<uint32_t>(padded + e2));
      hstage.push_back(static_cast<uint32_t>(raw + b));
      hstage.push_back(static_cast<uint32_t>(q) | (j + 1 == ns ? kStreamClosed : 0u));
    }
    for (int64_t j = 0; j < ns; j += kStreamWaves) {
      const int64_t cnt2 = ns - j < kStreamWaves ? ns - j : kStreamWaves;
      xcd_wgs[q % kXcds].push_back(static_cast<uint32_t>(slices + j));
      xcd_wgs[q % kXcds].push_back(static_cast<uint32_t>(cnt2));
    }
    slices += ns;
    padded += (len + kStreamTile - 1) / kStreamTile * kStreamTile;
  }
  if (hs[panels] != static_cast<uint32_t>(L.nnz_long) || slices > L.slices_cap ||
      padded > L.entries_cap)
    return hipErrorInvalidValue;
  // Work-group b runs on XCD b % 8: give it a work-group of a panel that XCD
  // stages (q % 8), so each XCD's panels stay in its L2; when an XCD's list
  // runs out, the longest remaining list fills in.
  std::vector<uint32_t> hwg;
  std::vector<size_t> pos(kXcds, 0);
  int64_t left = 0;
  for (const auto& v : xcd_wgs) left += static_cast<int64_t>(v.size() / 2);
  for (int64_t b = 0; left > 0; ++b) {
    int x = static_cast<int>(b % kXcds);
    if (pos[x] >= xcd_wgs[x].size()) {
      size_t best = 0;
      for (int y = 0; y < kXcds; ++y) {
        const size_t rem = xcd_wgs[y].size() - pos[y];
        if (rem > best) {
          best = rem;
          x = y;
        }
      }
    }
    hwg.push_back(xcd_wgs[x][pos[x]]);
    hwg.push_back(xcd_wgs[x][pos[x] + 1]);
    pos[x] += 2;
    --left;
  }
  uint32_t* dshift = small + panels + 2;
  e = hipMemcpyAsync(dshift, shift.data(), sizeof(uint32_t) * shift.size(),
                     hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !hstage.empty())
    e = hipMemcpyAsync(stage, hstage.data(), sizeof(uint32_t) * hstage.size(),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !hwg.empty())
    e = hipMemcpyAsync(wg_table, hwg.data(), sizeof(uint32_t) * hwg.size(),
                       hipMemcpyHostToDevice, s);
  // Zeroed stream (padding reads as column 0, value 0).
  if (e == hipSuccess) e = fill_launch(scol, 0, sizeof(uint16_t) * static_cast<size_t>(padded), s);
  if (e == hipSuccess) e = fill_launch(sval, 0, vb * static_cast<size_t>(padded), s);
  if (e != hipSuccess) return e;
#define ASC_SCATTER(I, VB)                                                                     \
  stream_scatter_kernel<I, VB><<<sb, kBlock, 0, s>>>(                                        \
      static_cast<const I*>(a.indices), static_cast<const VB*>(a.values), sorder, sstart,    \
      spanel, lens_s, lrow_s, dest, dshift, nseg, pp->panel_cols, scol,                     \
      static_cast<VB*>(sval), seg_lrow, at(L.seg_src), at(L.seg_dst), at(L.seg_len))
  if (a.index_bytes == 4) {
    if (vb == 8) ASC_SCATTER(int32_t, uint64_t);
    else if (vb == 4) ASC_SCATTER(int32_t, uint32_t);
    else ASC_SCATTER(int32_t, uint16_t);
  } else {
    if (vb == 8) ASC_SCATTER(int64_t, uint64_t);
    else if (vb == 4) ASC_SCATTER(int64_t, uint32_t);
    else ASC_SCATTER(int64_t, uint16_t);
  }
#undef ASC_SCATTER
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (slices > 0) {
    stream_slice_kernel<<<static_cast<unsigned>((slices + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        stage, slices, pseg, dest, scol, slice_info);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const int64_t centries = panels * (L.groups + 1);
  stream_ctable_kernel<<<static_cast<unsigned>((centries + kBlock - 1) / kBlock), kBlock, 0, s>>>(
      pseg, lrow_s, panels, L.groups, ctable);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // The host tables above must outlive their copies.
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  pp->slices = slices;
  pp->wgs = static_cast<int64_t>(hwg.size() / 2);
  pp->nseg = nseg;
  return hipSuccess;
}

hipError_t panel_stream_refresh_launch(const PanelPlan& pp, const void* values,
                                       size_t value_bytes, hipStream_t s) {
  if (!pp.stream || pp.nseg == 0) return hipSuccess;
  const unsigned sb = static_cast<unsigned>((pp.nseg + kBlock - 1) / kBlock);
#define ASC_REFRESH(VB)                                                                    \
  stream_refresh_kernel<VB><<<sb, kBlock, 0, s>>>(static_cast<const VB*>(values), pp.seg_src, \
                                                  pp.seg_dst, pp.seg_len, pp.nseg,          \
                                                  static_cast<VB*>(const_cast<void*>(pp.sval)))
  if (value_bytes == 8) ASC_REFRESH(uint64_t);
  else if (value_bytes == 4) ASC_REFRESH(uint32_t);
  else if (value_bytes == 2) ASC_REFRESH(uint16_t);
  else return hipErrorInvalidValue;
#undef ASC_REFRESH
  return hipGetLastError();
}

hipError_t spmv_launch(const SpmvArgs& a, void* ws, hipStream_t s, const SideStream* side) {
  if (a.a.outer == 0) return hipSuccess;
  KParams kp;
  kp.offsets = a.a.offsets;
  kp.indices = a.a.indices;
  kp.values = a.a.values;
  kp.rows = a.a.outer;
  kp.nnz = a.a.nnz;
  kp.x = a.x;
  kp.incx = a.incx;
  kp.y = a.y;
  kp.incy = a.incy;
  kp.alpha = a.alpha;
  kp.beta = a.beta;
  kp.form = static_cast<int>(a.form);
  {
    const size_t vb = elem_bytes(a.a.elem);
    kp.paired = vb <= 4 && a.a.index_bytes == 4 &&
                reinterpret_cast<uintptr_t>(a.a.values) % (2 * vb) == 0 &&
                reinterpret_cast<uintptr_t>(a.a.indices) % 8 == 0;
  }
  switch (a.a.elem) {
    case Elem::F64: return launch_by_index<double, false>(a, kp, ws, s, side);
    case Elem::F32:
      return a.twosum ? launch_by_index<float, true>(a, kp, ws, s, side)
                      : launch_by_index<float, false>(a, kp, ws, s, side);
    case Elem::BF16:
      return a.twosum ? launch_by_index<bf16_t, true>(a, kp, ws, s, side)
                      : launch_by_index<bf16_t, false>(a, kp, ws, s, side);
    case Elem::F16:
      return a.twosum ? launch_by_index<f16_t, true>(a, kp, ws, s, side)
                      : launch_by_index<f16_t, false>(a, kp, ws, s, side);
  }
  return hipErrorInvalidValue;
}

}  // namespace asc_hip
