// Sixteen sparsity features of a CSR matrix, computed on the GPU for the
// learned kernel selector.  One pass over the offsets and column indices;
// per-block partial sums (FP64) are reduced by a single block in fixed
// order, so the result is deterministic.
//
//  f0  log2(rows+1)              f8  frac rows, length in [4, 16)
//  f1  log2(cols+1)              f9  frac rows, length in [16, 64)
//  f2  log2(nnz+1)               f10 frac rows, length in [64, 256)
//  f3  log2(1+mean row length)   f11 frac rows, length >= 256
//  f4  row-length CV (std/mean)  f12 frac nnz in rows of length >= 256
//  f5  log2(1+max row length)    f13 mean |col - diag(row)| / cols
//  f6  frac empty rows           f14 frac nnz with |col - diag(row)| < 4096
//  f7  frac rows, length [1, 4)  f15 mean log2(1 + column gap) within rows
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kSums = 14;  // accumulated quantities (see below)
// sums: 0 len, 1 len^2, 2 empty, 3 [1,4), 4 [4,16), 5 [16,64), 6 [64,256),
//       7 >=256, 8 nnz in rows >=256, 9 sum |c-d|/cols, 10 near count,
//       11 sum log2(1+gap), 12 gap count, 13 max len

template <typename I>
__global__ __launch_bounds__(kBlock) void features_partial_kernel(
    const int64_t* __restrict__ offsets, const I* __restrict__ idx, int64_t rows,
    int64_t cols, double* __restrict__ partial) {
  __shared__ double s_red[kSums][kBlock / kWave];
  double acc[kSums];
#pragma unroll
  for (int k = 0; k < kSums; ++k) acc[k] = 0.0;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int64_t wave_id = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + wave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
  const double scale = rows > 0 ? static_cast<double>(cols) / static_cast<double>(rows) : 0.0;
  // One wave per row (grid-stride): lanes walk the row's columns.
  for (int64_t r = wave_id; r < rows; r += nwaves) {
    const int64_t b = offsets[r], e = offsets[r + 1];
    const int64_t len = e - b;
    const double diag = (static_cast<double>(r) + 0.5) * scale;
    double dist = 0.0, near = 0.0, lgap = 0.0, ngap = 0.0;
    for (int64_t k = b + lane; k < e; k += kWave) {
      const double c = static_cast<double>(idx[k]);
      const double d = c > diag ? c - diag : diag - c;
      dist += d;
      near += d < 4096.0 ? 1.0 : 0.0;
      if (k > b) {
        const double g = c - static_cast<double>(idx[k - 1]);
        lgap += log2(1.0 + g);
        ngap += 1.0;
      }
    }
    if (lane == 0) {
      const double l = static_cast<double>(len);
      acc[0] += l;
      acc[1] += l * l;
      acc[2] += len == 0 ? 1.0 : 0.0;
      acc[3] += (len >= 1 && len < 4) ? 1.0 : 0.0;
      acc[4] += (len >= 4 && len < 16) ? 1.0 : 0.0;
      acc[5] += (len >= 16 && len < 64) ? 1.0 : 0.0;
      acc[6] += (len >= 64 && len < 256) ? 1.0 : 0.0;
      acc[7] += len >= 256 ? 1.0 : 0.0;
      acc[8] += len >= 256 ? l : 0.0;
      acc[13] = l > acc[13] ? l : acc[13];
    }
    acc[9] += cols > 0 ? dist / static_cast<double>(cols) : 0.0;
    acc[10] += near;
    acc[11] += lgap;
    acc[12] += ngap;
  }
  // Wave reduction (max for slot 13), then across the block's waves.
#pragma unroll
  for (int k = 0; k < kSums; ++k) {
    double v = acc[k];
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off);
      v = k == 13 ? (o > v ? o : v) : v + o;
    }
    if (lane == 0) s_red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < kSums) {
    const int k = threadIdx.x;
    double v = s_red[k][0];
    for (int w = 1; w < kBlock / kWave; ++w)
      v = k == 13 ? (s_red[k][w] > v ? s_red[k][w] : v) : v + s_red[k][w];
    partial[static_cast<int64_t>(blockIdx.x) * kSums + k] = v;
  }
}

__global__ __launch_bounds__(kBlock) void features_final_kernel(
    const double* __restrict__ partial, int nblocks, int64_t rows, int64_t cols,
    int64_t nnz, float* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s[kSums];
  for (int k = 0; k < kSums; ++k) s[k] = 0.0;
  for (int b = 0; b < nblocks; ++b)
    for (int k = 0; k < kSums; ++k) {
      const double v = partial[static_cast<int64_t>(b) * kSums + k];
      s[k] = k == 13 ? (v > s[k] ? v : s[k]) : s[k] + v;
    }
  const double R = rows > 0 ? static_cast<double>(rows) : 1.0;
  const double N = nnz > 0 ? static_cast<double>(nnz) : 1.0;
  const double mean = s[0] / R;
  const double var = s[1] / R - mean * mean;
  out[0] = static_cast<float>(log2(static_cast<double>(rows) + 1.0));
  out[1] = static_cast<float>(log2(static_cast<double>(cols) + 1.0));
  out[2] = static_cast<float>(log2(static_cast<double>(nnz) + 1.0));
  out[3] = static_cast<float>(log2(1.0 + mean));
  out[4] = static_cast<float>(mean > 0.0 ? sqrt(var > 0.0 ? var : 0.0) / mean : 0.0);
  out[5] = static_cast<float>(log2(1.0 + s[13]));
  for (int k = 0; k < 6; ++k) out[6 + k] = static_cast<float>(s[2 + k] / R);
  out[12] = static_cast<float>(s[8] / N);
  out[13] = static_cast<float>(s[9] / N);
  out[14] = static_cast<float>(s[10] / N);
  out[15] = static_cast<float>(s[12] > 0.0 ? s[11] / s[12] : 0.0);
}

constexpr int kFeatureBlocks = 1024;

}  // namespace

size_t features_workspace_bytes() { return sizeof(double) * kSums * kFeatureBlocks; }

hipError_t features_launch(const Compressed& a, float* out16_device, void* ws,
                           hipStream_t s) {
  double* partial = static_cast<double*>(ws);
  int64_t blocks = (a.outer + (kBlock / kWave) - 1) / (kBlock / kWave);
  if (blocks > kFeatureBlocks) blocks = kFeatureBlocks;
  if (blocks < 1) blocks = 1;
  if (a.index_bytes == 4)
    features_partial_kernel<int32_t><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
        a.offsets, static_cast<const int32_t*>(a.indices), a.outer, a.inner, partial);
  else
    features_partial_kernel<int64_t><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
        a.offsets, static_cast<const int64_t*>(a.indices), a.outer, a.inner, partial);
  features_final_kernel<<<1, kBlock, 0, s>>>(partial, static_cast<int>(blocks), a.outer,
                                             a.inner, a.nnz, out16_device);
  return hipGetLastError();
}

}  // namespace asc_hip
