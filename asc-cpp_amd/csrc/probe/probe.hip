// Measurement probes for gfx950: the ceilings behind the SpMV rooflines.
// Built into their own library (libasc_sparse_probe.so), used by the
// benchmark and its tests; the Sparse BLAS library does not contain them.
//
// Gather probe: every SpMV nonzero streams its index and value and gathers
// one x element.  This kernel does exactly that over synthetic indices, with
// nothing else around it (no rows, no reductions per row), so its rate is
// the ceiling a random-column SpMV can reach when x of a given size is
// resident at a given cache level (LDS aside).  Indices are clamped to the x
// extent on the device, so any input stays in bounds.
//
// Row gather probe: the SpMM's B-row gathers without the SpMM around them.
// Every index gathers one 256-byte row (32 FP64) of a table, an 8-lane
// group per row and four rows in flight per group, as the plain SpMM kernel
// does; timed over random indices into tables of a given size it gives the
// rate a random-column SpMM can reach when B (or its pass slice) is
// resident at that cache level.
//
// Stream probes: a plain HBM read (every 16-byte word summed, one partial
// per work-group) and a plain copy, with nontemporal 16-byte loads, each
// work-group streaming one contiguous chunk (measured faster than a
// grid-stride walk: 6.8 against 6.2 TB/s for a 4 GiB read): the measured HBM
// ceilings beside the 8 TB/s peak.
#include "device_common.h"
#include "probe_launch.h"

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

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kRowProbeUnroll = 4;

__global__ __launch_bounds__(kBlock) void row_gather_probe_kernel(
    const int32_t* __restrict__ idx, int64_t n, const f64x4* __restrict__ table, uint32_t last,
    double* __restrict__ partials) {
  __shared__ double s_wave[kBlock / kWave];
  constexpr int U = kRowProbeUnroll;
  const int q = threadIdx.x & 7;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * (kBlock / 8);
  const int64_t g = static_cast<int64_t>(blockIdx.x) * (kBlock / 8) + (threadIdx.x >> 3);
  f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
  int64_t e0 = g * U;
  for (; e0 + U <= n; e0 += groups * U) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t c0 = static_cast<uint32_t>(ld_stream(idx + e0 + u));
      c[u] = c0 < last ? c0 : last;
    }
    f64x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = table[static_cast<int64_t>(c[u]) * 8 + q];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
  }
  for (int64_t e = e0; e < n && e < e0 + U; ++e) {  // the last partial batch
    const uint32_t c0 = static_cast<uint32_t>(ld_stream(idx + e));
    acc += table[static_cast<int64_t>(c0 < last ? c0 : last) * 8 + q];
  }
  double s = acc.x + acc.y + acc.z + acc.w;
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

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Work-group b's chunk of `words` 16-byte words.
__device__ __forceinline__ void chunk_of(int64_t words, int64_t* begin, int64_t* end) {
  const int64_t per = (words + gridDim.x - 1) / gridDim.x;
  const int64_t b = static_cast<int64_t>(blockIdx.x) * per;
  *begin = b < words ? b : words;
  *end = b + per < words ? b + per : words;
}

// Sums every 32-bit word of the work-group's chunk (wrapping), so no load is
// dead; eight loads in flight per lane.
__global__ __launch_bounds__(kBlock) void stream_read_kernel(const u32x4* __restrict__ src,
                                                             int64_t words,
                                                             uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_wave[kBlock / kWave];
  int64_t begin, end;
  chunk_of(words, &begin, &end);
  uint32_t s = 0;
  int64_t i = begin + threadIdx.x;
  for (; i + 7 * kBlock < end; i += 8 * kBlock) {
    u32x4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(src + i + u * kBlock);
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  for (; i < end; i += kBlock) {
    const u32x4 v = __builtin_nontemporal_load(src + i);
    s += v.x + v.y + v.z + v.w;
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x % kWave == 0) s_wave[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += s_wave[w];
    partials[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void stream_copy_kernel(const u32x4* __restrict__ src,
                                                             u32x4* __restrict__ dst,
                                                             int64_t words) {
  int64_t begin, end;
  chunk_of(words, &begin, &end);
  int64_t i = begin + threadIdx.x;
  for (; i + 3 * kBlock < end; i += 4 * kBlock) {
    u32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * kBlock);
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * kBlock);
  }
  for (; i < end; i += kBlock) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
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

hipError_t row_gather_probe_launch(const int32_t* idx, int64_t n, const double* table,
                                   int64_t table_rows, double* partials, hipStream_t s) {
  if (n == 0 || table_rows == 0) return hipSuccess;
  row_gather_probe_kernel<<<static_cast<unsigned>(kRowGatherProbeBlocks), kBlock, 0, s>>>(
      idx, n, reinterpret_cast<const f64x4*>(table), static_cast<uint32_t>(table_rows - 1),
      partials);
  return hipGetLastError();
}

hipError_t stream_probe_launch(const void* src, void* dst, int64_t words, uint32_t* partials,
                               hipStream_t s) {
  if (words == 0) return hipSuccess;
  const u32x4* in = static_cast<const u32x4*>(src);
  if (dst == nullptr)
    stream_read_kernel<<<static_cast<unsigned>(kStreamProbeBlocks), kBlock, 0, s>>>(in, words,
                                                                                   partials);
  else
    stream_copy_kernel<<<static_cast<unsigned>(kStreamProbeBlocks), kBlock, 0, s>>>(
        in, static_cast<u32x4*>(dst), words);
  return hipGetLastError();
}

}  // namespace asc_hip
