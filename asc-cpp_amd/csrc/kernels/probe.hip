// Gather-rate probe for gfx950: the measurement behind the SpMV rooflines.
//
// Every SpMV nonzero streams its index and value and gathers one x element.
// This kernel does exactly that over synthetic indices, with nothing else
// around it (no rows, no reductions per row), so its rate is the ceiling a
// random-column SpMV can reach when x of a given size is resident at a given
// cache level (LDS aside).  Indices are clamped to the x extent on the
// device, so any input stays in bounds.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void gather_probe_kernel(
    const int32_t* __restrict__ idx, const double* __restrict__ val, int64_t n,
    const double* __restrict__ x, uint32_t last, double* __restrict__ partials) {
  __shared__ double s_wave[kBlock / kWave];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  double s = 0.0;
  int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
#pragma unroll 8
  for (; i < n; i += stride) {
    const uint32_t c0 = static_cast<uint32_t>(ld_stream(idx + i));
    const uint32_t c = c0 < last ? c0 : last;
    s += ld_stream(val + i) * x[c];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x % kWave == 0) s_wave[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t += s_wave[w];
    partials[blockIdx.x] = t;
  }
}

}  // namespace

hipError_t gather_probe_launch(const int32_t* idx, const double* val, int64_t n,
                               const double* x, int64_t x_extent, double* partials,
                               hipStream_t s) {
  if (n == 0 || x_extent == 0) return hipSuccess;
  gather_probe_kernel<<<static_cast<unsigned>(kGatherProbeBlocks), kBlock, 0, s>>>(
      idx, val, n, x, static_cast<uint32_t>(x_extent - 1), partials);
  return hipGetLastError();
}

}  // namespace asc_hip
