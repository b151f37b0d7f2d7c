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
constexpr int kWideBatches = 4;

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
  // Column passes: splits[(q-1)*rows + r] is the position, relative to the
  // start of row r, of its first entry in pass q's column range.
  const uint32_t* splits;
  int64_t passes;
  int64_t pass;
};

// The row's entries in this pass; with one pass, the whole row.
__device__ __forceinline__ void pass_range(const MParams& p, int64_t row, int64_t* b,
                                           int64_t* e) {
  const int64_t rb = p.offsets[row];
  *b = p.pass == 0 ? rb : rb + p.splits[(p.pass - 1) * p.rows + row];
  *e = p.pass + 1 == p.passes ? p.offsets[row + 1] : rb + p.splits[p.pass * p.rows + row];
}

// Split table of the column passes: one thread per row, an advancing binary
// search per pass boundary over the row's sorted columns.
template <typename I>
__global__ __launch_bounds__(kBlock) void spmm_split_kernel(
    const int64_t* __restrict__ offsets, const I* __restrict__ col, int64_t rows,
    int64_t width, int64_t passes, uint32_t* __restrict__ splits) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (row >= rows) return;
  const int64_t b = offsets[row];
  const int64_t e = offsets[row + 1];
  int64_t lo = b;
  for (int64_t q = 1; q < passes; ++q) {
    const int64_t bound = q * width;
    int64_t hi = e;
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo) >> 1);
      if (static_cast<int64_t>(col[mid]) < bound)
        lo = mid + 1;
      else
        hi = mid;
    }
    splits[(q - 1) * rows + row] = static_cast<uint32_t>(lo - b);
  }
}

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
  int64_t begin, end;
  pass_range(p, row, &begin, &end);
  // Rows with no entry in this pass leave C untouched.  Passes run in column
  // order, so every C(i,j) still receives its terms in stored order.
  if (p.passes > 1 && begin == end) return;  // uniform over the G-group
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

// k <= 32: one wave per output row.  Lanes 0-31 hold C(i, j = lane) and take
// the row's even-numbered entries, lanes 32-63 the odd ones, so one wave
// instruction gathers two B rows.  Each batch's column indices and values
// are loaded while the previous batch's B rows are in flight.  The odd
// lanes' products move to the even lanes by a cross-lane shuffle and every
// C(i,j) adds its terms strictly in stored order: the result is the serial
// reference's, bit for bit.
template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmm_pair_kernel(MParams p) {
#pragma clang fp contract(off)
  constexpr int kRowsPerBlock = kBlock / kWave;
  constexpr int U = kUnroll;  // entries per lane per batch (2U per wave)
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = blk * kRowsPerBlock + threadIdx.x / kWave;
  if (row >= p.rows) return;  // whole wave
  const int lane = threadIdx.x % kWave;
  const int slot = lane >> 5;
  const int64_t j = lane & 31;
  const bool jvalid = j < p.k;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ b = static_cast<const T*>(p.b);
  T* c = static_cast<T*>(p.c);
  const T alpha = static_cast<T>(p.alpha);
  int64_t begin, end;
  pass_range(p, row, &begin, &end);
  if (p.passes > 1 && begin == end) return;  // uniform over the wave
  const int64_t cpos = p.c_row_major ? row * p.ldc + j : j * p.ldc + row;
  T acc = (slot == 0 && jvalid) ? c[cpos] : T(0);
  int64_t cur_c[U];
  T cur_v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t e = begin + 2 * u + slot;
    cur_c[u] = e < end ? static_cast<int64_t>(col[e]) : 0;
    cur_v[u] = e < end ? alpha * val[e] : T(0);
  }
  for (int64_t e0 = begin; e0 < end; e0 += 2 * U) {
    T bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = e0 + 2 * u + slot;
      bv[u] = (e < end && jvalid)
                  ? (p.b_row_major ? b[cur_c[u] * p.ldb + j] : b[j * p.ldb + cur_c[u]])
                  : T(0);
    }
    int64_t nxt_c[U];
    T nxt_v[U];
    const int64_t e1 = e0 + 2 * U;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = e1 + 2 * u + slot;
      nxt_c[u] = e < end ? static_cast<int64_t>(col[e]) : 0;
      nxt_v[u] = e < end ? alpha * val[e] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T term = cur_v[u] * bv[u];
      const T odd = __shfl(term, static_cast<int>(j) + 32);
      if (e0 + 2 * u < end) acc = acc + term;      // uniform over the wave
      if (e0 + 2 * u + 1 < end) acc = acc + odd;  // uniform over the wave
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cur_c[u] = nxt_c[u];
      cur_v[u] = nxt_v[u];
    }
  }
  if (slot == 0 && jvalid) c[cpos] = acc;
}

// Row-major B, k <= 32 and a multiple of 4, 4-element aligned rows: one wave
// per output row, eight entries per instruction.  Lane l takes entry slot
// l / 8 and columns 4*(l % 8) .. +3, loading them as one 4-wide vector, so a
// wave instruction gathers eight B rows.  kWideBatches batches of eight are
// in flight, with the next batches' indices and values loaded meanwhile.
// Each batch's products go through LDS to the slot-0 lanes, which own
// C(i, 4q..4q+3) and add the terms in stored order: bit-exact with the
// serial reference.  Waves of a block run different rows, so they
// synchronise only within the wave.
template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmm_wide_kernel(MParams p) {
#pragma clang fp contract(off)
  typedef T v4 __attribute__((ext_vector_type(4)));
  constexpr int kRowsPerBlock = kBlock / kWave;
  constexpr int U = kWideBatches;
  __shared__ T s_terms[kRowsPerBlock][8][32];
  const int w = threadIdx.x / kWave;
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t row = blk * kRowsPerBlock + w;
  if (row >= p.rows) return;  // whole wave
  const int lane = threadIdx.x % kWave;
  const int g = lane >> 3;
  const int64_t jq = 4 * (lane & 7);
  const bool qvalid = jq < p.k;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ b = static_cast<const T*>(p.b);
  T* c = static_cast<T*>(p.c);
  const T alpha = static_cast<T>(p.alpha);
  int64_t begin, end;
  pass_range(p, row, &begin, &end);
  if (p.passes > 1 && begin == end) return;  // uniform over the wave
  const bool owner = g == 0 && qvalid;
  auto cpos = [&](int64_t j) { return p.c_row_major ? row * p.ldc + j : j * p.ldc + row; };
  T acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = owner ? c[cpos(jq + i)] : T(0);
  int64_t cur_c[U];
  T cur_v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t e = begin + 8 * u + g;
    cur_c[u] = e < end ? static_cast<int64_t>(col[e]) : 0;
    cur_v[u] = e < end ? alpha * val[e] : T(0);
  }
  for (int64_t e0 = begin; e0 < end; e0 += 8 * U) {
    v4 bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = e0 + 8 * u + g;
      if (e < end && qvalid)
        bv[u] = *reinterpret_cast<const v4*>(b + cur_c[u] * p.ldb + jq);
      else
        bv[u] = v4{T(0), T(0), T(0), T(0)};
    }
    int64_t nxt_c[U];
    T nxt_v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = e0 + 8 * U + 8 * u + g;
      nxt_c[u] = e < end ? static_cast<int64_t>(col[e]) : 0;
      nxt_v[u] = e < end ? alpha * val[e] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T t0 = cur_v[u] * bv[u].x;
      const T t1 = cur_v[u] * bv[u].y;
      const T t2 = cur_v[u] * bv[u].z;
      const T t3 = cur_v[u] * bv[u].w;
      T* slot = &s_terms[w][g][jq];
      slot[0] = t0;
      slot[1] = t1;
      slot[2] = t2;
      slot[3] = t3;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (owner) {
        const int64_t eb = e0 + 8 * u;
#pragma unroll
        for (int sl = 0; sl < 8; ++sl) {
          if (eb + sl < end) {
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] = acc[i] + s_terms[w][sl][jq + i];
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cur_c[u] = nxt_c[u];
      cur_v[u] = nxt_v[u];
    }
  }
  if (owner) {
#pragma unroll
    for (int i = 0; i < 4; ++i) c[cpos(jq + i)] = acc[i];
  }
}

template <typename T, typename I>
hipError_t launch_typed(const MParams& p, hipStream_t s) {
  const uintptr_t vec_bytes = 4 * sizeof(T);
  if (p.k <= 32 && p.k % 4 == 0 && p.b_row_major && p.ldb % 4 == 0 &&
      reinterpret_cast<uintptr_t>(p.b) % vec_bytes == 0) {
    constexpr int rows_per_block = kBlock / kWave;
    const unsigned grid =
        static_cast<unsigned>((p.rows + rows_per_block - 1) / rows_per_block);
    spmm_wide_kernel<T, I><<<grid, kBlock, 0, s>>>(p);
    return hipGetLastError();
  }
  if (p.k <= 32) {
    constexpr int rows_per_block = kBlock / kWave;
    const unsigned grid =
        static_cast<unsigned>((p.rows + rows_per_block - 1) / rows_per_block);
    spmm_pair_kernel<T, I><<<grid, kBlock, 0, s>>>(p);
    return hipGetLastError();
  }
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

namespace {

// Column passes and the bytes of their split table (0 for one pass).
int64_t pass_width(const SpmmArgs& a) {
  const int64_t inner = a.a.inner;
  if (a.panel_cols <= 0 || a.panel_cols >= inner) return inner;
  if (a.a.nnz > static_cast<int64_t>(UINT32_MAX)) return inner;  // 32-bit split positions
  return a.panel_cols;
}

int64_t pass_count(const SpmmArgs& a) {
  const int64_t w = pass_width(a);
  return w > 0 && w < a.a.inner ? (a.a.inner - 1) / w + 1 : 1;
}

}  // namespace

size_t spmm_workspace_bytes(const SpmmArgs& a) {
  const int64_t passes = pass_count(a);
  if (passes <= 1) return 0;
  return (sizeof(uint32_t) * static_cast<size_t>(passes - 1) * static_cast<size_t>(a.a.outer) +
          255) & ~size_t(255);
}

hipError_t spmm_launch(const SpmmArgs& a, void* workspace, hipStream_t s) {
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
  p.passes = pass_count(a);
  p.splits = static_cast<const uint32_t*>(workspace);
  const bool i32 = a.a.index_bytes == 4;
  if (p.passes > 1) {
    const unsigned blocks = static_cast<unsigned>((p.rows + kBlock - 1) / kBlock);
    uint32_t* splits = static_cast<uint32_t*>(workspace);
    if (i32)
      spmm_split_kernel<int32_t><<<blocks, kBlock, 0, s>>>(
          p.offsets, static_cast<const int32_t*>(p.indices), p.rows, pass_width(a), p.passes, splits);
    else
      spmm_split_kernel<int64_t><<<blocks, kBlock, 0, s>>>(
          p.offsets, static_cast<const int64_t*>(p.indices), p.rows, pass_width(a), p.passes, splits);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int64_t q = 0; q < p.passes; ++q) {
    p.pass = q;
    hipError_t e = hipErrorInvalidValue;
    if (a.a.elem == Elem::F64)
      e = i32 ? launch_typed<double, int32_t>(p, s) : launch_typed<double, int64_t>(p, s);
    else if (a.a.elem == Elem::F32)
      e = i32 ? launch_typed<float, int32_t>(p, s) : launch_typed<float, int64_t>(p, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace asc_hip
