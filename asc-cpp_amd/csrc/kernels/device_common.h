// Device helpers shared by the gfx950 kernels: value storage types, wave64
// reductions, streaming loads and the XCD-aware block remap.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asc_hip {

constexpr int kWave = 64;   // CDNA wavefront width
constexpr int kXcds = 8;    // MI355X: 8 XCDs, blocks dispatched round-robin

// ---- storage-only reduced precisions -------------------------------------
// BF16 and FP16 values are stored as raw 16-bit words; arithmetic is FP32.
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

__device__ __forceinline__ float to_float(bf16_t v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}
__device__ __forceinline__ float to_float(f16_t v) {
  _Float16 h;
  __builtin_memcpy(&h, &v.bits, 2);
  return static_cast<float>(h);
}
__device__ __forceinline__ float to_float(float v) { return v; }
__device__ __forceinline__ double to_double(double v) { return v; }

// Accumulator type for each stored value type.
template <typename V> struct acc_of { using type = float; };
template <> struct acc_of<double> { using type = double; };

template <typename A, typename V>
__device__ __forceinline__ A widen(V v) { return static_cast<A>(to_float(v)); }
template <>
__device__ __forceinline__ double widen<double, double>(double v) { return v; }

// ---- streaming loads -----------------------------------------------------
// Matrix arrays are read exactly once per SpMV: non-temporal loads keep them
// from evicting the gathered x vector from L2.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ bf16_t ld_stream(const bf16_t* p) {
  bf16_t r;
  r.bits = __builtin_nontemporal_load(&p->bits);
  return r;
}
__device__ __forceinline__ f16_t ld_stream(const f16_t* p) {
  f16_t r;
  r.bits = __builtin_nontemporal_load(&p->bits);
  return r;
}

// Row-walking kernels with few lanes per row leave a cache line of the
// matrix arrays shared by the row groups of neighbouring waves; a
// non-temporal load does not keep it for the second reader, so they read
// the matrix with ordinary cached loads (measured on banded and stencil
// matrices: up to 2x).  From 16 lanes per row a wave reads whole lines, and
// the non-temporal loads keep gathered x in L2.
template <bool NT, typename T>
__device__ __forceinline__ T ld_matrix(const T* p) {
  if constexpr (NT) return ld_stream(p);
  else return *p;
}
template <bool NT, typename I>
__device__ __forceinline__ int64_t load_index_as(const I* p, int64_t k) {
  return static_cast<int64_t>(ld_matrix<NT>(p + k));
}

// ---- wave reductions -----------------------------------------------------
// Butterfly sum over aligned groups of G lanes (G a power of two <= 64).
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// ---- XCD-aware block remap -----------------------------------------------
// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// receives a contiguous range of logical tiles (bijective for any count),
// which keeps neighbouring tiles' shared x/B lines in one L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t block, int64_t nblocks) {
  if (nblocks < kXcds) return block;
  const int64_t xcd = block % kXcds;
  const int64_t local = block / kXcds;
  const int64_t q = nblocks / kXcds, r = nblocks % kXcds;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

__device__ __forceinline__ int64_t load_index(const int32_t* p, int64_t k) {
  return static_cast<int64_t>(ld_stream(p + k));
}
__device__ __forceinline__ int64_t load_index(const int64_t* p, int64_t k) {
  return ld_stream(p + k);
}

}  // namespace asc_hip
