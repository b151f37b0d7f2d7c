// Small element-wise kernels: byte fill, index narrowing and value-type
// conversion with round-to-nearest-even.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void fill_kernel(uint8_t* dst, uint8_t byte,
                                                      size_t bytes) {
  const uint32_t word = 0x01010101u * byte;
  const size_t words = bytes / 4;
  const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
  size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 3) == 0;
  if (aligned) {
    uint32_t* d32 = reinterpret_cast<uint32_t*>(dst);
    for (size_t w = i; w < words; w += stride) d32[w] = word;
    for (size_t t = words * 4 + i; t < bytes; t += stride) dst[t] = byte;
  } else {
    for (size_t t = i; t < bytes; t += stride) dst[t] = byte;
  }
}

__global__ __launch_bounds__(kBlock) void narrow_kernel(
    const int64_t* __restrict__ src, int32_t* __restrict__ dst, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
       i < n; i += stride)
    dst[i] = static_cast<int32_t>(ld_stream(src + i));
}

// FP64 -> FP32 rounded to odd: truncate, then set the last bit if inexact.
// Rounding to odd at 24 bits followed by round-to-nearest-even at 8 or 11
// bits equals a single correct rounding (the intermediate has >= 2 extra
// bits), which avoids double-rounding errors.
__device__ __forceinline__ float round_to_odd_f32(double d) {
  float f = __double2float_rz(d);
  if (static_cast<double>(f) != d && d == d) {
    uint32_t bits = __float_as_uint(f) | 1u;
    f = __uint_as_float(bits);
  }
  return f;
}

__device__ __forceinline__ uint16_t to_bf16_rne(double d) {
  if (d != d) return 0x7FC0;  // quiet NaN
  const float f = round_to_odd_f32(d);
  const uint32_t bits = __float_as_uint(f);
  const uint32_t rounded = bits + 0x7FFFu + ((bits >> 16) & 1u);
  return static_cast<uint16_t>(rounded >> 16);
}

__device__ __forceinline__ uint16_t to_f16_rne(double d) {
  const _Float16 h = static_cast<_Float16>(round_to_odd_f32(d));
  uint16_t bits;
  __builtin_memcpy(&bits, &h, 2);
  return bits;
}

__global__ __launch_bounds__(kBlock) void convert_kernel(
    const double* __restrict__ src, void* __restrict__ dst, int to, int64_t n,
    double scale) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
       i < n; i += stride) {
    const double d = ld_stream(src + i) * scale;
    if (to == static_cast<int>(Elem::F32)) {
      static_cast<float*>(dst)[i] = static_cast<float>(d);
    } else if (to == static_cast<int>(Elem::BF16)) {
      static_cast<uint16_t*>(dst)[i] = to_bf16_rne(d);
    } else {
      static_cast<uint16_t*>(dst)[i] = to_f16_rne(d);
    }
  }
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace

hipError_t fill_launch(void* dst, uint8_t byte, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  fill_kernel<<<grid_for(static_cast<int64_t>(bytes / 4 + 1)), kBlock, 0, s>>>(
      static_cast<uint8_t*>(dst), byte, bytes);
  return hipGetLastError();
}

hipError_t narrow_indices_launch(const int64_t* src, int32_t* dst, int64_t n,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  narrow_kernel<<<grid_for(n), kBlock, 0, s>>>(src, dst, n);
  return hipGetLastError();
}

hipError_t convert_values_launch(const double* src, void* dst, Elem to,
                                 int64_t n, double scale, hipStream_t s) {
  if (n == 0) return hipSuccess;
  convert_kernel<<<grid_for(n), kBlock, 0, s>>>(src, dst, static_cast<int>(to), n, scale);
  return hipGetLastError();
}

}  // namespace asc_hip
