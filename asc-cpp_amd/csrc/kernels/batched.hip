// Batched compat SpMV for gfx950: thousands of independent CSR matrices
// y_m = alpha_m*A_m*x_m + beta_m*y_m in one persistent launch.
//
// The host builds a work list: every item is a contiguous row range of one
// matrix plus the lanes-per-row width G chosen for that matrix's size class
// (average row length bucket).  Items hold roughly the same number of
// nonzeros, so a fixed grid of persistent workgroups walks contiguous item
// ranges (XCD-remapped) and the launch is independent of the batch count.
// No counters or atomics: the kernel replays identically inside a hipGraph.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;

template <typename T, typename I>
__device__ __forceinline__ void run_item(const BatchMatrix& m, int64_t r0, int64_t r1,
                                         int g) {
  const T* __restrict__ val = static_cast<const T*>(m.values);
  const I* __restrict__ col = static_cast<const I*>(m.indices);
  const T* __restrict__ x = static_cast<const T*>(m.x);
  T* y = static_cast<T*>(m.y);
  const int lane = threadIdx.x & (g - 1);
  const int rows_per_pass = kBlock / g;
  const T alpha = static_cast<T>(m.alpha), beta = static_cast<T>(m.beta);
  for (int64_t row = r0 + threadIdx.x / g; row < r1; row += rows_per_pass) {
    const int64_t b = m.offsets[row], e = m.offsets[row + 1];
    T s0 = T(0), s1 = T(0);
    int64_t k = b + lane;
    for (; k + g < e; k += 2 * g) {
      const int64_t c0 = static_cast<int64_t>(col[k]), c1 = static_cast<int64_t>(col[k + g]);
      s0 += val[k] * x[c0 * m.incx];
      s1 += val[k + g] * x[c1 * m.incx];
    }
    if (k < e) s0 += val[k] * x[static_cast<int64_t>(col[k]) * m.incx];
    T s = s0 + s1;
    for (int off = g >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, g);
    if (lane == 0) {
      T r = alpha * s;
      if (beta != T(0)) r += beta * y[row * m.incy];
      y[row * m.incy] = r;
    }
  }
}

// CSR-stream item (Greathouse & Daga, SC'14): the work-group multiplies the
// item's entries by x in one coalesced sweep into LDS, then each thread sums
// its rows' products from LDS.  Short rows then stream at full width instead
// of leaving most lanes of a row group idle.  An item with more entries than
// the LDS window takes the lane path instead (returns false).
template <typename T, typename I>
__device__ __forceinline__ bool run_stream_item(const BatchMatrix& m, int64_t r0, int64_t r1,
                                                T* s_prod) {
  const T* __restrict__ val = static_cast<const T*>(m.values);
  const I* __restrict__ col = static_cast<const I*>(m.indices);
  const T* __restrict__ x = static_cast<const T*>(m.x);
  T* y = static_cast<T*>(m.y);
  const int64_t b0 = m.offsets[r0];
  const int64_t n = m.offsets[r1] - b0;
  if (n > kStreamEntries) return false;  // uniform over the work-group
  for (int64_t t = threadIdx.x; t < n; t += kBlock)
    s_prod[t] = val[b0 + t] * x[static_cast<int64_t>(col[b0 + t]) * m.incx];
  __syncthreads();
  const T alpha = static_cast<T>(m.alpha), beta = static_cast<T>(m.beta);
  for (int64_t row = r0 + threadIdx.x; row < r1; row += kBlock) {
    const int64_t b = m.offsets[row] - b0, e = m.offsets[row + 1] - b0;
    T sum = T(0);
    for (int64_t p = b; p < e; ++p) sum += s_prod[p];
    T r = alpha * sum;
    if (beta != T(0)) r += beta * y[row * m.incy];
    y[row * m.incy] = r;
  }
  __syncthreads();  // the window is reused by the next item
  return true;
}

// Work-group w walks the item range the host balanced by estimated entries
// (XCD-remapped, so neighbouring ranges, often of one matrix, share an L2).
template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void batched_spmv_kernel(
    const BatchMatrix* __restrict__ mats, const BatchItem* __restrict__ items,
    const int64_t* __restrict__ ranges) {
  __shared__ T s_prod[kStreamEntries];
  const int64_t wg = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i0 = ranges[wg];
  const int64_t i1 = ranges[wg + 1];
  for (int64_t i = i0; i < i1; ++i) {
    const BatchItem it = items[i];
    const BatchMatrix& m = mats[it.matrix];
    if (it.lanes == 0 && run_stream_item<T, I>(m, it.row_begin, it.row_end, s_prod)) continue;
    run_item<T, I>(m, it.row_begin, it.row_end, it.lanes == 0 ? 1 : it.lanes);
  }
}

}  // namespace

hipError_t batched_spmv_launch(const BatchLaunch& L, hipStream_t s) {
  if (L.nitems == 0) return hipSuccess;
  const unsigned grid = static_cast<unsigned>(L.grid);
  if (L.elem == Elem::F64) {
    if (L.index_bytes == 4)
      batched_spmv_kernel<double, int32_t><<<grid, kBlock, 0, s>>>(L.mats, L.items, L.ranges);
    else
      batched_spmv_kernel<double, int64_t><<<grid, kBlock, 0, s>>>(L.mats, L.items, L.ranges);
  } else {
    if (L.index_bytes == 4)
      batched_spmv_kernel<float, int32_t><<<grid, kBlock, 0, s>>>(L.mats, L.items, L.ranges);
    else
      batched_spmv_kernel<float, int64_t><<<grid, kBlock, 0, s>>>(L.mats, L.items, L.ranges);
  }
  return hipGetLastError();
}

}  // namespace asc_hip
