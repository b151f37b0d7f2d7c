// Standardized SpMM  C += alpha*op(A)*B  for gfx950 (blas.h:1692-1750).
//
// The ABI hands this kernel a compressed operand whose outer dimension is
// the output row (op(A) already in row form).  G lanes share one output
// row, lane l owning output columns l, l+G, ...; every lane walks the row's
// stored entries in order and updates its C(i,j) with
//     C(i,j) = C(i,j) + (alpha*v)*B(c,j)
// unfused, which is the serial reference's per-entry arithmetic
// (blas.h:1736-1748), so the result is bit-exact with it.  For the
// row-major k=32 block a G=32 group reads each gathered B row as one
// contiguous 256-byte segment and writes C rows the same way.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 8;

struct MParams {
  const int64_t* offsets;
  const void* indices;
  const void* values;
  int64_t rows;
  const void* b;
  int64_t ldb;
  int b_row_major;
  void* c;
  int64_t ldc;
  int c_row_major;
  int64_t k;
  double alpha;
};

template <int G, typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmm_kernel(MParams p) {
#pragma clang fp contract(off)
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = blk * (kBlock / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (row >= p.rows) return;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ b = static_cast<const T*>(p.b);
  T* c = static_cast<T*>(p.c);
  const T alpha = static_cast<T>(p.alpha);
  const int64_t begin = p.offsets[row];
  const int64_t end = p.offsets[row + 1];
  for (int64_t j = lane; j < p.k; j += G) {
    const int64_t cpos = p.c_row_major ? row * p.ldc + j : j * p.ldc + row;
    T acc = c[cpos];
    int64_t e = begin;
    // kUnroll B gathers in flight per lane; the additions stay in stored
    // order, so the result is unchanged.
    for (; e + kUnroll <= end; e += kUnroll) {
      T av[kUnroll], bv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t src = static_cast<int64_t>(col[e + u]);
        av[u] = alpha * val[e + u];
        bv[u] = p.b_row_major ? b[src * p.ldb + j] : b[j * p.ldb + src];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const T term = av[u] * bv[u];
        acc = acc + term;
      }
    }
    for (; e < end; ++e) {
      const int64_t src = static_cast<int64_t>(col[e]);
      const T av = alpha * val[e];
      const T bv = p.b_row_major ? b[src * p.ldb + j] : b[j * p.ldb + src];
      const T term = av * bv;
      acc = acc + term;
    }
    c[cpos] = acc;
  }
}

template <typename T, typename I>
hipError_t launch_typed(const MParams& p, hipStream_t s) {
  int g = 1;
  while (g < p.k && g < 64) g <<= 1;
  auto grid = [&](int G) {
    return static_cast<unsigned>((p.rows + (kBlock / G) - 1) / (kBlock / G));
  };
  switch (g) {
    case 1: spmm_kernel<1, T, I><<<grid(1), kBlock, 0, s>>>(p); break;
    case 2: spmm_kernel<2, T, I><<<grid(2), kBlock, 0, s>>>(p); break;
    case 4: spmm_kernel<4, T, I><<<grid(4), kBlock, 0, s>>>(p); break;
    case 8: spmm_kernel<8, T, I><<<grid(8), kBlock, 0, s>>>(p); break;
    case 16: spmm_kernel<16, T, I><<<grid(16), kBlock, 0, s>>>(p); break;
    case 32: spmm_kernel<32, T, I><<<grid(32), kBlock, 0, s>>>(p); break;
    default: spmm_kernel<64, T, I><<<grid(64), kBlock, 0, s>>>(p); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t spmm_launch(const SpmmArgs& a, hipStream_t s) {
  if (a.a.outer == 0 || a.k == 0) return hipSuccess;
  MParams p;
  p.offsets = a.a.offsets;
  p.indices = a.a.indices;
  p.values = a.a.values;
  p.rows = a.a.outer;
  p.b = a.b;
  p.ldb = a.ldb;
  p.b_row_major = a.b_row_major;
  p.c = a.c;
  p.ldc = a.ldc;
  p.c_row_major = a.c_row_major;
  p.k = a.k;
  p.alpha = a.alpha;
  const bool i32 = a.a.index_bytes == 4;
  if (a.a.elem == Elem::F64)
    return i32 ? launch_typed<double, int32_t>(p, s) : launch_typed<double, int64_t>(p, s);
  if (a.a.elem == Elem::F32)
    return i32 ? launch_typed<float, int32_t>(p, s) : launch_typed<float, int64_t>(p, s);
  return hipErrorInvalidValue;
}

}  // namespace asc_hip
