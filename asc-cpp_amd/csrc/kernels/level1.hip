// Sparse Level-1 kernels for gfx950 (blas.h:1330-1582) over a canonical
// indexed vector (strictly increasing indices, so every dense element is
// touched at most once) and a strided dense vector (element i at x[i*inc]).
//
//   dot (plain)    one 1024-thread block: thread t sums entries t, t+1024, ...
//                  in order, then a fixed-shape tree: deterministic run to run.
//   dot (ordered)  one lane, left to right from +0 without FMA: the serial
//                  reference's arithmetic bit for bit.
//   axpy / gather / gather-zero / scatter   one thread per entry; each dense
//                  element is updated once, so these are exact.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kDotBlock = 1024;

template <typename T>
__global__ __launch_bounds__(kDotBlock) void dot_plain_kernel(
    const int64_t* __restrict__ idx, const T* __restrict__ v, const T* __restrict__ x,
    int64_t inc, int64_t nnz, T* __restrict__ out) {
  __shared__ T s_part[kDotBlock];
  T sum = T(0);
  for (int64_t p = threadIdx.x; p < nnz; p += kDotBlock) sum += v[p] * x[idx[p] * inc];
  s_part[threadIdx.x] = sum;
  __syncthreads();
  for (int half = kDotBlock / 2; half > 0; half >>= 1) {
    if (static_cast<int>(threadIdx.x) < half) s_part[threadIdx.x] += s_part[threadIdx.x + half];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s_part[0];
}

// Lane u's value (u uniform over the wave), for every lane.
__device__ __forceinline__ float lane_value(float t, int u) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), u));
}
__device__ __forceinline__ double lane_value(double t, int u) {
  const int64_t b = __double_as_longlong(t);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), u));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), u));
  return __longlong_as_double(static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo));
}

template <typename T>
__global__ __launch_bounds__(kWave) void dot_ordered_kernel(const int64_t* __restrict__ idx,
                                                            const T* __restrict__ v,
                                                            const T* __restrict__ x, int64_t inc,
                                                            int64_t nnz, T* __restrict__ out) {
#pragma clang fp contract(off)
  // One wave: its lanes form 64 terms v[p]*x[idx[p]] at once (the next 64
  // loading meanwhile), and the running sum takes them one after another in
  // entry order, each read from its lane: the serial sum, bit for bit.
  const int lane = static_cast<int>(threadIdx.x);
  const auto term_of = [&](int64_t p) { return p < nnz ? v[p] * x[idx[p] * inc] : T(0); };
  T r = T(0);
  T term = term_of(lane);
  for (int64_t p0 = 0; p0 < nnz; p0 += kWave) {
    const T cur = term;
    term = term_of(p0 + kWave + lane);
    if (nnz - p0 >= kWave) {
#pragma unroll
      for (int u = 0; u < kWave; ++u) r = r + lane_value(cur, u);
    } else {
      const int n = static_cast<int>(nnz - p0);
      for (int u = 0; u < n; ++u) r = r + lane_value(cur, u);
    }
  }
  if (lane == 0) *out = r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void level1_kernel(int op, T alpha,
                                                       const int64_t* __restrict__ idx,
                                                       T* __restrict__ v, T* __restrict__ x,
                                                       int64_t inc, int64_t nnz) {
#pragma clang fp contract(off)
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; p < nnz; p += stride) {
    T* xp = x + idx[p] * inc;
    switch (op) {
      case 1: {  // axpy
        const T scaled = alpha * v[p];
        *xp = *xp + scaled;
        break;
      }
      case 2: v[p] = *xp; break;  // gather
      case 3:                     // gather-zero
        v[p] = *xp;
        *xp = T(0);
        break;
      default: *xp = v[p]; break;  // scatter
    }
  }
}

template <typename T>
hipError_t level1_typed(int op, bool ordered, double alpha, const int64_t* idx, void* v, void* x,
                        int64_t inc, int64_t nnz, void* result, hipStream_t s) {
  T* tv = static_cast<T*>(v);
  T* tx = static_cast<T*>(x);
  if (op == 0) {
    if (ordered)
      dot_ordered_kernel<T><<<1, 64, 0, s>>>(idx, tv, tx, inc, nnz, static_cast<T*>(result));
    else
      dot_plain_kernel<T><<<1, kDotBlock, 0, s>>>(idx, tv, tx, inc, nnz, static_cast<T*>(result));
    return hipGetLastError();
  }
  if (nnz == 0) return hipSuccess;
  int64_t blocks = (nnz + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  level1_kernel<T><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(op, static_cast<T>(alpha), idx,
                                                                     tv, tx, inc, nnz);
  return hipGetLastError();
}

}  // namespace

hipError_t level1_launch(int op, bool ordered, Elem elem, double alpha, const int64_t* indices,
                         void* values, void* dense, int64_t inc, int64_t nnz, void* result,
                         hipStream_t s) {
  if (elem == Elem::F64)
    return level1_typed<double>(op, ordered, alpha, indices, values, dense, inc, nnz, result, s);
  return level1_typed<float>(op, ordered, alpha, indices, values, dense, inc, nnz, result, s);
}

}  // namespace asc_hip
