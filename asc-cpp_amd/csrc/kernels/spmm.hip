// Standardized SpMM  C += alpha*op(A)*B  for gfx950 (blas.h:1692-1750).
//
// The ABI hands these kernels a compressed operand whose outer dimension is
// the output row (op(A) already in row form).  Two forms:
//  - ordered: G lanes share one output row, lane l owning output columns
//    l, l+G, ...; every lane walks the row's stored entries in order and
//    updates its C(i,j) with
//        C(i,j) = C(i,j) + (alpha*v)*B(c,j)
//    unfused, which is the serial reference's per-entry arithmetic
//    (blas.h:1736-1748), so the result is bit-exact with it;
//  - plain: per-row register accumulation with fused multiply-adds and one
//    C update per row (below), within the plain bound of the serial result.
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
  // Rows of at least this many entries are left to the panel path.
  int64_t skip_min;
  int64_t inner;  // columns of op(A) (rows of B)
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
  // Rows a column-panel plan lists run segment by segment elsewhere.
  if (p.offsets[row + 1] - p.offsets[row] >= p.skip_min) return;
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

// ---- plain mode (any association of each C(i,j)'s terms) ------------------
//
// Row-major B and C, k <= 32 with k, ldb and ldc multiples of 4 and B and C
// aligned to 4 elements.  An 8-lane group covers one output row: lane q owns
// columns 4q..4q+3 and gathers them as one 4-wide vector per stored entry,
// so a group reads a whole B row (k = 32) per instruction.  Every lane
// accumulates in registers with fused multiply-adds and adds alpha times its
// sums to C at the end.
//
// A block takes kBlock consecutive rows and sorts them, in LDS, into short
// rows (fewer than kPlainLongRow entries in the pass) and long rows:
//   short: one 8-lane group per row, kPlainUnroll B rows in flight per group;
//   long:  one wave per row, its eight groups taking every eighth entry
//          (kPlainWideBatches batches of eight B rows in flight) and summing
//          their partial sums by cross-lane exchange at the end.
// Which group or wave takes a row depends on LDS atomics, but the arithmetic
// of a row does not, so results repeat bit for bit from run to run.
// Rows with no entries in the pass leave C untouched.
constexpr int kPlainLongRow = 64;
constexpr int kPlainUnroll = 4;
constexpr int kPlainWideBatches = 4;

template <typename T>
using vec4_of = T __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ void fma4(vec4_of<T>& acc, T v, const vec4_of<T>& b) {
  acc.x = __builtin_fma(v, b.x, acc.x);
  acc.y = __builtin_fma(v, b.y, acc.y);
  acc.z = __builtin_fma(v, b.z, acc.z);
  acc.w = __builtin_fma(v, b.w, acc.w);
}

template <typename T>
__device__ __forceinline__ void plain_store(const MParams& p, int64_t row, int64_t jq,
                                           vec4_of<T> acc) {
  typedef vec4_of<T> v4;
  T* c = static_cast<T*>(p.c) + row * p.ldc + jq;
  const T alpha = static_cast<T>(p.alpha);
  v4 old = *reinterpret_cast<const v4*>(c);
  fma4<T>(old, alpha, acc);
  *reinterpret_cast<v4*>(c) = old;
}

template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void spmm_plain_kernel(MParams p) {
  typedef vec4_of<T> v4;
  __shared__ int64_t s_begin[kBlock];
  __shared__ int64_t s_end[kBlock];
  __shared__ uint16_t s_short[kBlock];
  __shared__ uint16_t s_long[kBlock];
  __shared__ uint32_t s_count[2];
  const int tid = threadIdx.x;
  const int64_t row0 = xcd_remap(blockIdx.x, gridDim.x) * kBlock;
  if (tid < 2) s_count[tid] = 0;
  __syncthreads();
  if (row0 + tid < p.rows &&
      p.offsets[row0 + tid + 1] - p.offsets[row0 + tid] < p.skip_min) {
    int64_t begin, end;
    pass_range(p, row0 + tid, &begin, &end);
    s_begin[tid] = begin;
    s_end[tid] = end;
    if (end - begin >= kPlainLongRow)
      s_long[atomicAdd(&s_count[1], 1u)] = static_cast<uint16_t>(tid);
    else if (end > begin)
      s_short[atomicAdd(&s_count[0], 1u)] = static_cast<uint16_t>(tid);
  }
  __syncthreads();
  const int q = tid & 7;
  const int64_t jq = 4 * q;
  const bool qvalid = jq < p.k;
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ b = static_cast<const T*>(p.b) + jq;
  const v4 zero = v4{T(0), T(0), T(0), T(0)};

  // Long rows: wave w takes the block's long rows w, w + 4, ...
  const uint32_t nlong = s_count[1];
  const int g = (tid % kWave) >> 3;  // entry slot within the wave
  for (uint32_t i = tid / kWave; i < nlong; i += kBlock / kWave) {
    const int r = s_long[i];
    const int64_t begin = s_begin[r], end = s_end[r];
    v4 acc = zero;
    constexpr int U = kPlainWideBatches;
    for (int64_t e0 = begin; e0 < end; e0 += 8 * U) {
      int64_t c[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t e = e0 + 8 * u + g;
        c[u] = e < end ? static_cast<int64_t>(col[e]) : 0;
        v[u] = e < end ? val[e] : T(0);
      }
      v4 bv[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        bv[u] = (e0 + 8 * u + g < end && qvalid) ? *reinterpret_cast<const v4*>(b + c[u] * p.ldb)
                                                 : zero;
#pragma unroll
      for (int u = 0; u < U; ++u) fma4<T>(acc, v[u], bv[u]);
    }
    // The eight slots' sums of the same columns, in a fixed tree.
#pragma unroll
    for (int m = 8; m < kWave; m <<= 1) {
      acc.x += __shfl_xor(acc.x, m);
      acc.y += __shfl_xor(acc.y, m);
      acc.z += __shfl_xor(acc.z, m);
      acc.w += __shfl_xor(acc.w, m);
    }
    if (g == 0 && qvalid) plain_store<T>(p, row0 + r, jq, acc);
  }

  // Short rows: group gi takes the block's short rows gi, gi + 32, ...
  const uint32_t nshort = s_count[0];
  if (!qvalid) return;
  for (uint32_t i = tid >> 3; i < nshort; i += kBlock / 8) {
    const int r = s_short[i];
    const int64_t begin = s_begin[r], end = s_end[r];
    v4 acc = zero;
    constexpr int U = kPlainUnroll;
    for (int64_t e0 = begin; e0 < end; e0 += U) {
      int64_t c[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = e0 + u < end;
        c[u] = in ? static_cast<int64_t>(col[e0 + u]) : 0;
        v[u] = in ? val[e0 + u] : T(0);
      }
      v4 bv[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        bv[u] = e0 + u < end ? *reinterpret_cast<const v4*>(b + c[u] * p.ldb) : zero;
#pragma unroll
      for (int u = 0; u < U; ++u) fma4<T>(acc, v[u], bv[u]);
    }
    plain_store<T>(p, row0 + r, jq, acc);
  }
}

// ---- plain form over a column-panel plan (listed rows) ------------------------
// The plan cuts op(A)'s columns into panels; a listed row's entries in panel
// q form one segment, contiguous in the CSR arrays (seg_src, seg_len), and
// segments are numbered panel by panel (ctable gives each panel's range).
// Per slice of W (16 or 32) columns of B and C:
//   spmm_seg_kernel: a 16-lane group takes one segment at a time, lane j
//     gathering columns j0 + (W/16)*j .. of each entry's B row (a whole
//     256-byte row for k = 32, so a panel's slice of B is contiguous), FMA
//     accumulation in stored order, one partial row per segment.  Work-group
//     w runs on XCD w mod 8 and takes a chunk of panel q (q mod 8 = w mod 8,
//     panels in turn), so an XCD's work-groups gather from one panel's B
//     slice at a time (16384 rows x 32 FP64 columns = 4 MiB at the stream
//     analysis's widest FP64 panel), which its L2 keeps;
//   spmm_panel_combine_kernel: per group of kStreamGroupRows listed rows,
//     16 of the slice's columns and part of the panels, the panels' partial
//     rows added in panel order into LDS accumulators (a row has at most one
//     segment per panel, so the threads of one panel never share an
//     accumulator);
//   spmm_panel_finish_kernel: C(r, j) += alpha * (sum of the parts in order).
// The order of every sum is fixed by the plan, so results repeat bit for bit.
constexpr int kSegLanes = 16;  // lanes a segment (one DPP row)
constexpr int kSegStep = 16;  // entries a group gathers at once
constexpr int kSegBlock = 256;
constexpr int kCombineBlock = 1024;
// The combine splits each group's panels into `parts` ranges (one
// work-group each, at most kCombineMaxParts), chosen so that the work-groups
// (groups x column halves x parts, one 1024-thread work-group per CU for its
// 128 KiB of LDS) come to about two per CU of an MI355X: on the 640M-entry
// Zipf matrix (50 groups, 100 units) 4 parts 1.20 ms, 5 parts 1.00 ms,
// 8 parts 1.22 ms, 16 parts 1.10 ms.
constexpr int kCombineMaxParts = 16;
constexpr int64_t kCombineTargetWgs = 2 * 256;
// Chunks per panel: about one segment per 16-lane group (16 segments a
// 256-thread work-group), so that an XCD's work-groups are short-lived and
// dispatched in panel order, and the XCD's L2 holds one panel's slice of B
// at a time.  Measured on the 640M-entry Zipf matrix (k = 32, 16384-column
// panels, rows of 4096+ entries listed): 256 segments a work-group (at most
// 256 chunks a panel) 8.10 ms, 98 segments 7.97 ms, 16 segments 7.65 ms; 8
// segments over 128 threads 7.77 ms, 4 over 64 threads 7.82 ms.
constexpr int64_t kSegChunksMax = int64_t(1) << 20;
constexpr int64_t kSegChunkMinSegs = kSegBlock / kSegLanes;

// Lane u's value, for every lane of its 16-lane row (DPP row_newbcast:u).
template <int U>
__device__ __forceinline__ int row_bcast(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + U, 0xF, 0xF, false);
}
template <int U>
__device__ __forceinline__ double row_bcast(double v) {
  const int64_t b = __double_as_longlong(v);
  const int lo = row_bcast<U>(static_cast<int>(b));
  const int hi = row_bcast<U>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) |
                              static_cast<int64_t>(static_cast<uint32_t>(lo)));
}
template <int U>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(row_bcast<U>(__float_as_int(v)));
}

// Column pair a lane gathers and accumulates (k > 16: two adjacent columns
// in one vector load), or one column (k <= 16).
template <typename T, int CPL> struct ColPack;
template <typename T> struct ColPack<T, 1> {
  T x;
  __device__ static ColPack load(const char* p) { return {*reinterpret_cast<const T*>(p)}; }
  __device__ void fma(T v, const ColPack& b) { x = __builtin_fma(v, b.x, x); }
};
template <typename T> struct ColPack<T, 2> {
  T x, y;
  __device__ static ColPack load(const char* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 t = *reinterpret_cast<const v2*>(p);
    return {t.x, t.y};
  }
  __device__ void fma(T v, const ColPack& b) {
    x = __builtin_fma(v, b.x, x);
    y = __builtin_fma(v, b.y, y);
  }
};

// One step of a segment: the kSegStep entries from t (lane j of the 16-lane
// group loaded entry t + j%kSegStep into c, v; c < 0 past the segment's end).  Their B rows
// are gathered together (byte offset c * row_bytes from the lane's columns,
// 32-bit when B spans less than 4 GiB), then accumulated in stored order.
template <typename T, int CPL, bool SMALL, int U>
__device__ __forceinline__ void seg_gather(const char* __restrict__ bj, uint32_t row_bytes, int c,
                                           bool on, ColPack<T, CPL> (&bv)[kSegStep]) {
  if constexpr (U < kSegStep) {
    // Every lane takes part in the broadcast (lane U holds entry U); only
    // lanes with columns inside k gather.
    const int cu = row_bcast<U>(c);
    if (on && cu >= 0) {
      if constexpr (SMALL)
        bv[U] = ColPack<T, CPL>::load(bj + __umul24(static_cast<uint32_t>(cu), row_bytes));
      else
        bv[U] = ColPack<T, CPL>::load(bj + static_cast<int64_t>(cu) * row_bytes);
    } else {
      bv[U] = ColPack<T, CPL>{};
    }
    seg_gather<T, CPL, SMALL, U + 1>(bj, row_bytes, c, on, bv);
  }
}
template <typename T, int CPL, int U>
__device__ __forceinline__ void seg_accumulate(T v, const ColPack<T, CPL> (&bv)[kSegStep],
                                               ColPack<T, CPL>& acc) {
  if constexpr (U < kSegStep) {
    acc.fma(row_bcast<U>(v), bv[U]);
    seg_accumulate<T, CPL, U + 1>(v, bv, acc);
  }
}

template <typename T, typename I, int CPL, bool SMALL>
__global__ __launch_bounds__(kSegBlock) void spmm_seg_kernel(
    const I* __restrict__ col, const T* __restrict__ val, const T* __restrict__ b, int64_t ldb,
    int64_t j0, int64_t kk, const uint32_t* __restrict__ seg_src,
    const uint32_t* __restrict__ seg_len, const uint32_t* __restrict__ ctable, int64_t groups,
    int64_t panels, int64_t chunks, T* __restrict__ partial) {
  constexpr int G = kSegLanes;  // lanes a segment
  constexpr int W = G * CPL;   // columns a segment's partial row
  const int64_t w = blockIdx.x;
  const int64_t step = w / kXcds;
  const int64_t q = w % kXcds + kXcds * (step / chunks);
  const int64_t m = step % chunks;
  if (q >= panels) return;
  const uint64_t o0 = ctable[q * (groups + 1)], o1 = ctable[q * (groups + 1) + groups];
  const uint64_t a = o0 + (o1 - o0) * m / chunks;
  const uint64_t e = o0 + (o1 - o0) * (m + 1) / chunks;
  const int j = threadIdx.x % G;
  const int jr = j % kSegStep;  // the entry this lane loads in a step
  const bool on = CPL * j < kk;  // kk is a multiple of CPL
  const char* __restrict__ bj = reinterpret_cast<const char*>(b + j0 + CPL * j);
  const uint32_t row_bytes = static_cast<uint32_t>(ldb * static_cast<int64_t>(sizeof(T)));
  for (uint64_t o = a + threadIdx.x / G; o < e; o += kSegBlock / G) {
    const int64_t src = ld_stream(seg_src + o);
    const int64_t len = ld_stream(seg_len + o);
    // Lane j holds entry t + j%kSegStep of the step (c = -1 past the end);
    // the next step's entries load while this step's B rows are gathered.
    int c = -1;
    T v = T(0);
    if (jr < len) {
      c = static_cast<int>(ld_stream(col + src + jr));
      v = ld_stream(val + src + jr);
    }
    ColPack<T, CPL> acc{};
    for (int64_t t = 0; t < len; t += kSegStep) {
      ColPack<T, CPL> bv[kSegStep];
      seg_gather<T, CPL, SMALL, 0>(bj, row_bytes, c, on, bv);
      const T vcur = v;
      const int64_t tn = t + kSegStep + jr;
      c = -1;
      v = T(0);
      if (tn < len) {
        c = static_cast<int>(ld_stream(col + src + tn));
        v = ld_stream(val + src + tn);
      }
      seg_accumulate<T, CPL, 0>(vcur, bv, acc);
    }
    if (on) {
      T* out = partial + o * W + CPL * j;
      __builtin_nontemporal_store(acc.x, out);
      if constexpr (CPL == 2) __builtin_nontemporal_store(acc.y, out + 1);
    }
  }
}

// Ordered form over a column-panel plan: one launch per panel, in panel
// order.  A 16-lane group takes a segment of the panel, loads its row's C
// columns, and continues each C(r, j)'s serial chain C = C + (alpha*v)*B(c, j)
// (no FMA) over the segment's entries in stored order; a row's segments run
// in panel order, which is its stored order, so every chain is the serial
// reference's, bit for bit.  The segment's row: its group by binary search
// over the panel's combine table, then its tag.
template <typename T, int CPL>
__device__ __forceinline__ void ordered_add(ColPack<T, CPL>& acc, T av, const ColPack<T, CPL>& b) {
#pragma clang fp contract(off)
  acc.x = acc.x + av * b.x;
  if constexpr (CPL == 2) acc.y = acc.y + av * b.y;
}

// The step's first n entries added in stored order (every lane takes part in
// each broadcast; the adds stop at the segment's end).
template <typename T, int CPL, int U>
__device__ __forceinline__ void seg_ordered_accumulate(T av, const ColPack<T, CPL> (&bv)[kSegStep],
                                                       ColPack<T, CPL>& acc, int64_t n) {
  if constexpr (U < kSegStep) {
    const T avu = row_bcast<U>(av);
    if (U < n) ordered_add<T, CPL>(acc, avu, bv[U]);
    seg_ordered_accumulate<T, CPL, U + 1>(av, bv, acc, n);
  }
}

template <typename T, typename I, int CPL, bool SMALL>
__global__ __launch_bounds__(kSegBlock) void spmm_seg_ordered_kernel(
    const I* __restrict__ col, const T* __restrict__ val, const T* __restrict__ b, int64_t ldb,
    T* __restrict__ c, int64_t ldc, int64_t j0, int64_t kk, double alpha,
    const uint32_t* __restrict__ seg_src, const uint32_t* __restrict__ seg_len,
    const uint16_t* __restrict__ seg_lrow, const uint32_t* __restrict__ ctable,
    const uint32_t* __restrict__ long_rows, int64_t groups, int64_t q) {
#pragma clang fp contract(off)
  constexpr int G = kSegLanes;
  const uint32_t* __restrict__ ct = ctable + q * (groups + 1);
  const uint64_t o0 = ct[0], o1 = ct[groups];
  const int j = threadIdx.x % G;
  const bool on = CPL * j < kk;  // kk is a multiple of CPL
  const T a = static_cast<T>(alpha);
  const char* __restrict__ bj = reinterpret_cast<const char*>(b + j0 + CPL * j);
  const uint32_t row_bytes = static_cast<uint32_t>(ldb * static_cast<int64_t>(sizeof(T)));
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * (kSegBlock / G);
  for (uint64_t o = o0 + static_cast<uint64_t>(blockIdx.x) * (kSegBlock / G) + threadIdx.x / G;
       o < o1; o += stride) {
    int64_t lo = 0, hi = groups;  // the last group whose first ordinal is at most o
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) / 2;
      if (ct[mid] <= o) lo = mid;
      else hi = mid;
    }
    const int64_t row = long_rows[lo * kStreamGroupRows + seg_lrow[o]];
    T* cr = c + row * ldc + j0 + CPL * j;
    ColPack<T, CPL> acc{};
    if (on) {
      acc.x = cr[0];
      if constexpr (CPL == 2) acc.y = cr[1];
    }
    const int64_t src = seg_src[o];
    const int64_t len = seg_len[o];
    int cc = -1;
    T av = T(0);
    if (j < len) {
      cc = static_cast<int>(ld_stream(col + src + j));
      av = a * ld_stream(val + src + j);
    }
    for (int64_t t = 0; t < len; t += kSegStep) {
      ColPack<T, CPL> bv[kSegStep];
      seg_gather<T, CPL, SMALL, 0>(bj, row_bytes, cc, on, bv);
      const T avcur = av;
      const int64_t tn = t + kSegStep + j;
      cc = -1;
      av = T(0);
      if (tn < len) {
        cc = static_cast<int>(ld_stream(col + src + tn));
        av = a * ld_stream(val + src + tn);
      }
      seg_ordered_accumulate<T, CPL, 0>(avcur, bv, acc, len - t);
    }
    if (on) {
      cr[0] = acc.x;
      if constexpr (CPL == 2) cr[1] = acc.y;
    }
  }
}

// Ordered form, row by row: a 16-lane group per row (of the pass's entries
// when B is gathered in column passes), the same steps and serial chains as
// the segment kernel above.  Rows of at least skip_min entries are left to a
// column-panel plan.
template <typename T, typename I, int CPL, bool SMALL>
__global__ __launch_bounds__(kSegBlock) void spmm_rows_ordered_kernel(MParams p, int64_t j0,
                                                                      int64_t kk) {
#pragma clang fp contract(off)
  constexpr int G = kSegLanes;
  const int64_t row = (xcd_remap(blockIdx.x, gridDim.x) * kSegBlock + threadIdx.x) / G;
  if (row >= p.rows) return;  // whole group
  if (p.offsets[row + 1] - p.offsets[row] >= p.skip_min) return;
  int64_t begin, end;
  pass_range(p, row, &begin, &end);
  if (begin == end) return;
  const int j = threadIdx.x % G;
  const bool on = CPL * j < kk;  // kk is a multiple of CPL
  const T a = static_cast<T>(p.alpha);
  const I* __restrict__ col = static_cast<const I*>(p.indices);
  const T* __restrict__ val = static_cast<const T*>(p.values);
  const char* __restrict__ bj =
      reinterpret_cast<const char*>(static_cast<const T*>(p.b) + j0 + CPL * j);
  const uint32_t row_bytes = static_cast<uint32_t>(p.ldb * static_cast<int64_t>(sizeof(T)));
  T* cr = static_cast<T*>(p.c) + row * p.ldc + j0 + CPL * j;
  ColPack<T, CPL> acc{};
  if (on) {
    acc.x = cr[0];
    if constexpr (CPL == 2) acc.y = cr[1];
  }
  const int64_t len = end - begin;
  int cc = -1;
  T av = T(0);
  if (j < len) {
    cc = static_cast<int>(ld_stream(col + begin + j));
    av = a * ld_stream(val + begin + j);
  }
  for (int64_t t = 0; t < len; t += kSegStep) {
    ColPack<T, CPL> bv[kSegStep];
    seg_gather<T, CPL, SMALL, 0>(bj, row_bytes, cc, on, bv);
    const T avcur = av;
    const int64_t tn = t + kSegStep + j;
    cc = -1;
    av = T(0);
    if (tn < len) {
      cc = static_cast<int>(ld_stream(col + begin + tn));
      av = a * ld_stream(val + begin + tn);
    }
    seg_ordered_accumulate<T, CPL, 0>(avcur, bv, acc, len - t);
  }
  if (on) {
    cr[0] = acc.x;
    if constexpr (CPL == 2) cr[1] = acc.y;
  }
}

// Columns a combine work-group accumulates (kStreamGroupRows x 16 FP64 =
// 128 KiB of LDS): a 32-column partial row is combined in two halves.
constexpr int kCombineCols = 16;

// Panel ranges per combine group for `units` (groups x column halves).
int64_t combine_parts(int64_t units) {
  const int64_t p = (kCombineTargetWgs + units / 2) / (units > 0 ? units : 1);
  return p < 1 ? 1 : p > kCombineMaxParts ? kCombineMaxParts : p;
}

template <typename T, int W>
__global__ __launch_bounds__(kCombineBlock) void spmm_panel_combine_kernel(
    const T* __restrict__ partial, const uint16_t* __restrict__ seg_lrow,
    const uint32_t* __restrict__ ctable, int64_t groups, int64_t panels, int64_t nlong,
    int64_t nparts, T* __restrict__ parts) {
  constexpr int H = W / kCombineCols;  // column halves
  __shared__ T acc[kStreamGroupRows * kCombineCols];
  const int64_t part = blockIdx.x % nparts;
  const int64_t gh = blockIdx.x / nparts;
  const int64_t g = gh / H, jc = (gh % H) * kCombineCols;
  for (int i = threadIdx.x; i < kStreamGroupRows * kCombineCols; i += kCombineBlock) acc[i] = T(0);
  __syncthreads();
  const int64_t qa = panels * part / nparts, qb = panels * (part + 1) / nparts;
  for (int64_t q = qa; q < qb; ++q) {
    const uint64_t lo = ctable[q * (groups + 1) + g], hi = ctable[q * (groups + 1) + g + 1];
    for (uint64_t i = lo * kCombineCols + threadIdx.x; i < hi * kCombineCols; i += kCombineBlock) {
      const uint64_t o = i / kCombineCols;
      const int jj = static_cast<int>(i % kCombineCols);
      acc[seg_lrow[o] * kCombineCols + jj] += partial[o * W + jc + jj];
    }
    __syncthreads();  // the next panel may hold the same rows
  }
  const int64_t row0 = g * kStreamGroupRows;
  const int64_t rows = nlong - row0 < kStreamGroupRows ? nlong - row0 : kStreamGroupRows;
  T* out = parts + (part * nlong + row0) * W + jc;
  for (int64_t i = threadIdx.x; i < rows * kCombineCols; i += kCombineBlock)
    out[(i / kCombineCols) * W + i % kCombineCols] = acc[i];
}

template <typename T, int W>
__global__ __launch_bounds__(kBlock) void spmm_panel_finish_kernel(
    const T* __restrict__ parts, int64_t nlong, int64_t nparts,
    const uint32_t* __restrict__ long_rows, T* __restrict__ c, int64_t ldc, int64_t j0, int64_t kk,
    double alpha) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nlong * W) return;
  const int64_t li = i / W, j = i % W;
  if (j >= kk) return;
  T sum = T(0);
  for (int64_t part = 0; part < nparts; ++part) sum += parts[(part * nlong + li) * W + j];
  T* cp = c + static_cast<int64_t>(long_rows[li]) * ldc + j0 + j;
  *cp = __builtin_fma(static_cast<T>(alpha), sum, *cp);
}

// Columns a segment's partial row covers: all of k up to 32 (whole B rows
// for k = 32, so each panel's slice of B is contiguous), two a lane; 16 for
// k <= 16.
int64_t seg_cols(int64_t k) { return k > 16 ? 32 : 16; }

template <typename T, typename I>
hipError_t launch_panel_rows(const SpmmArgs& a, const PanelPlan& pp, char* ws, hipStream_t s) {
  const int64_t width = seg_cols(a.k);
  T* partial = reinterpret_cast<T*>(ws);
  T* parts = reinterpret_cast<T*>(
      ws + ((sizeof(T) * width * static_cast<size_t>(pp.nseg) + 255) & ~size_t(255)));
  const int64_t groups = pp.groups;
  int64_t chunks = pp.nseg / (pp.panels * kSegChunkMinSegs);
  chunks = chunks < 1 ? 1 : chunks > kSegChunksMax ? kSegChunksMax : chunks;
  const unsigned seg_grid =
      static_cast<unsigned>((pp.panels + kXcds - 1) / kXcds * kXcds * chunks);
  // Combine work-groups: each group's column halves times the panel parts.
  const int64_t nparts = combine_parts(groups * (width / kCombineCols));
  const unsigned comb_grid = static_cast<unsigned>(groups * (width / kCombineCols) * nparts);
  const unsigned fin_grid = static_cast<unsigned>((pp.nlong * width + kBlock - 1) / kBlock);
  // 32-bit B offsets (24-bit row index and row pitch) when they reach all of B.
  const int64_t row_bytes = a.ldb * static_cast<int64_t>(sizeof(T));
  const bool small = a.a.inner < (int64_t(1) << 24) && row_bytes < (int64_t(1) << 24) &&
                     a.a.inner * row_bytes < (int64_t(1) << 32);
  for (int64_t j0 = 0; j0 < a.k; j0 += width) {
    const int64_t kk = a.k - j0 < width ? a.k - j0 : width;
#define ASC_SEG(CPL, SMALL)                                                                   \
  spmm_seg_kernel<T, I, CPL, SMALL><<<seg_grid, kSegBlock, 0, s>>>(                           \
      static_cast<const I*>(a.a.indices), static_cast<const T*>(a.a.values),                 \
      static_cast<const T*>(a.b), a.ldb, j0, kk, pp.seg_src, pp.seg_len, pp.ctable, groups,  \
      pp.panels, chunks, partial)
    if (width == 32) {
      if (small) ASC_SEG(2, true); else ASC_SEG(2, false);
    } else {
      if (small) ASC_SEG(1, true); else ASC_SEG(1, false);
    }
#undef ASC_SEG
    if (width == 32) {
      spmm_panel_combine_kernel<T, 32><<<comb_grid, kCombineBlock, 0, s>>>(
          partial, pp.seg_lrow, pp.ctable, groups, pp.panels, pp.nlong, nparts, parts);
      spmm_panel_finish_kernel<T, 32><<<fin_grid, kBlock, 0, s>>>(
          parts, pp.nlong, nparts, pp.long_rows, static_cast<T*>(a.c), a.ldc, j0, kk, a.alpha);
    } else {
      spmm_panel_combine_kernel<T, 16><<<comb_grid, kCombineBlock, 0, s>>>(
          partial, pp.seg_lrow, pp.ctable, groups, pp.panels, pp.nlong, nparts, parts);
      spmm_panel_finish_kernel<T, 16><<<fin_grid, kBlock, 0, s>>>(
          parts, pp.nlong, nparts, pp.long_rows, static_cast<T*>(a.c), a.ldc, j0, kk, a.alpha);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Work-groups of one ordered panel launch (each 16-lane group walks the
// panel's segments at this stride).
constexpr unsigned kSegOrderedGrid = 1024;

// The ordered form's listed rows: one launch per panel, in panel order (a
// row's later segments continue its chains), no workspace.
template <typename T, typename I>
hipError_t launch_panel_ordered(const SpmmArgs& a, const PanelPlan& pp, hipStream_t s) {
  const int64_t width = seg_cols(a.k);
  const int64_t row_bytes = a.ldb * static_cast<int64_t>(sizeof(T));
  const bool small = a.a.inner < (int64_t(1) << 24) && row_bytes < (int64_t(1) << 24) &&
                     a.a.inner * row_bytes < (int64_t(1) << 32);
  for (int64_t j0 = 0; j0 < a.k; j0 += width) {
    const int64_t kk = a.k - j0 < width ? a.k - j0 : width;
    for (int64_t q = 0; q < pp.panels; ++q) {
#define ASC_OSEG(CPL, SMALL)                                                                   \
  spmm_seg_ordered_kernel<T, I, CPL, SMALL><<<kSegOrderedGrid, kSegBlock, 0, s>>>(             \
      static_cast<const I*>(a.a.indices), static_cast<const T*>(a.a.values),                  \
      static_cast<const T*>(a.b), a.ldb, static_cast<T*>(a.c), a.ldc, j0, kk, a.alpha,        \
      pp.seg_src, pp.seg_len, pp.seg_lrow, pp.ctable, pp.long_rows, pp.groups, q)
      if (width == 32) {
        if (small) ASC_OSEG(2, true); else ASC_OSEG(2, false);
      } else {
        if (small) ASC_OSEG(1, true); else ASC_OSEG(1, false);
      }
#undef ASC_OSEG
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// The row-by-row ordered kernel (plain-form operands: row-major B and C,
// k <= 32 in one slice of 16 or 32 columns).
template <typename T, typename I>
hipError_t launch_rows_ordered(const MParams& p, hipStream_t s) {
  const int64_t width = p.k > 16 ? 32 : 16;
  const int64_t row_bytes = p.ldb * static_cast<int64_t>(sizeof(T));
  const int64_t inner = p.inner;
  const bool small = inner < (int64_t(1) << 24) && row_bytes < (int64_t(1) << 24) &&
                     inner * row_bytes < (int64_t(1) << 32);
  const unsigned grid = static_cast<unsigned>((p.rows + kSegBlock / kSegLanes - 1) /
                                              (kSegBlock / kSegLanes));
  for (int64_t j0 = 0; j0 < p.k; j0 += width) {
    const int64_t kk = p.k - j0 < width ? p.k - j0 : width;
    if (width == 32) {
      if (small) spmm_rows_ordered_kernel<T, I, 2, true><<<grid, kSegBlock, 0, s>>>(p, j0, kk);
      else spmm_rows_ordered_kernel<T, I, 2, false><<<grid, kSegBlock, 0, s>>>(p, j0, kk);
    } else {
      if (small) spmm_rows_ordered_kernel<T, I, 1, true><<<grid, kSegBlock, 0, s>>>(p, j0, kk);
      else spmm_rows_ordered_kernel<T, I, 1, false><<<grid, kSegBlock, 0, s>>>(p, j0, kk);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T, typename I>
hipError_t launch_plain(const MParams& p, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((p.rows + kBlock - 1) / kBlock);
  spmm_plain_kernel<T, I><<<grid, kBlock, 0, s>>>(p);
  return hipGetLastError();
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

bool spmm_plain_applies(const SpmmArgs& a) {
  const size_t esz = a.a.elem == Elem::F64 ? 8 : 4;
  const uintptr_t vb = 4 * esz;
  return a.b_row_major && a.c_row_major && a.k <= 32 && a.k % 4 == 0 && a.ldb % 4 == 0 &&
         a.ldc % 4 == 0 && reinterpret_cast<uintptr_t>(a.b) % vb == 0 &&
         reinterpret_cast<uintptr_t>(a.c) % vb == 0;
}

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

size_t align_up(size_t n) { return (n + 255) & ~size_t(255); }

size_t split_bytes(const SpmmArgs& a) {
  const int64_t passes = pass_count(a);
  return passes <= 1 ? 0
                     : align_up(sizeof(uint32_t) * static_cast<size_t>(passes - 1) *
                                static_cast<size_t>(a.a.outer));
}

// Workspace of the panel path: one partial row per segment, then the
// combine's per-part row sums (parts x listed rows).
size_t panel_bytes(const SpmmArgs& a) {
  if (a.panel == nullptr || a.ordered) return 0;  // the ordered form needs none
  const size_t esz = a.a.elem == Elem::F64 ? 8 : 4;
  const int64_t width = seg_cols(a.k);
  const size_t nparts = static_cast<size_t>(combine_parts(a.panel->groups * (width / kCombineCols)));
  return align_up(esz * width * static_cast<size_t>(a.panel->nseg)) +
         align_up(esz * width * nparts * static_cast<size_t>(a.panel->nlong));
}

}  // namespace

size_t spmm_workspace_bytes(const SpmmArgs& a) { return split_bytes(a) + panel_bytes(a); }

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
  p.pass = 0;
  p.splits = static_cast<const uint32_t*>(workspace);
  const bool panel = a.panel != nullptr && !a.ordered && spmm_plain_applies(a);
  const bool opanel = a.panel != nullptr && a.ordered && spmm_plain_applies(a);
  p.skip_min = panel || opanel ? a.panel->long_min : INT64_MAX;
  p.inner = a.a.inner;
  const bool i32 = a.a.index_bytes == 4;
  const bool f64 = a.a.elem == Elem::F64;
  if (!f64 && a.a.elem != Elem::F32) return hipErrorInvalidValue;
  const bool plain = !a.ordered && spmm_plain_applies(a);
  // The ordered form on the plain form's operands runs row by row with 16
  // lanes a row (the same chains as the wide kernel).
  const bool rows_ordered = a.ordered && spmm_plain_applies(a);
  // The row-by-row kernels over every row (or, with a panel plan, the rows
  // it does not list).
  const auto run_rows = [&](hipStream_t st) -> hipError_t {
    hipError_t e;
    if (p.passes > 1) {
      const unsigned blocks = static_cast<unsigned>((p.rows + kBlock - 1) / kBlock);
      uint32_t* splits = static_cast<uint32_t*>(workspace);
      if (i32)
        spmm_split_kernel<int32_t><<<blocks, kBlock, 0, st>>>(
            p.offsets, static_cast<const int32_t*>(p.indices), p.rows, pass_width(a), p.passes,
            splits);
      else
        spmm_split_kernel<int64_t><<<blocks, kBlock, 0, st>>>(
            p.offsets, static_cast<const int64_t*>(p.indices), p.rows, pass_width(a), p.passes,
            splits);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int64_t q = 0; q < p.passes; ++q) {
      p.pass = q;
      if (plain)
        e = f64 ? (i32 ? launch_plain<double, int32_t>(p, st) : launch_plain<double, int64_t>(p, st))
                : (i32 ? launch_plain<float, int32_t>(p, st) : launch_plain<float, int64_t>(p, st));
      else if (rows_ordered)
        e = f64 ? (i32 ? launch_rows_ordered<double, int32_t>(p, st)
                       : launch_rows_ordered<double, int64_t>(p, st))
                : (i32 ? launch_rows_ordered<float, int32_t>(p, st)
                       : launch_rows_ordered<float, int64_t>(p, st));
      else
        e = f64 ? (i32 ? launch_typed<double, int32_t>(p, st) : launch_typed<double, int64_t>(p, st))
                : (i32 ? launch_typed<float, int32_t>(p, st) : launch_typed<float, int64_t>(p, st));
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  if (opanel && a.panel->nlong > 0) {
    const hipError_t e =
        f64 ? (i32 ? launch_panel_ordered<double, int32_t>(a, *a.panel, s)
                   : launch_panel_ordered<double, int64_t>(a, *a.panel, s))
            : (i32 ? launch_panel_ordered<float, int32_t>(a, *a.panel, s)
                   : launch_panel_ordered<float, int64_t>(a, *a.panel, s));
    if (e != hipSuccess) return e;
    return run_rows(s);
  }
  if (!panel || a.panel->nlong == 0) return run_rows(s);
  char* pws = static_cast<char*>(workspace) + split_bytes(a);
  const auto run_listed = [&](hipStream_t st) -> hipError_t {
    return f64 ? (i32 ? launch_panel_rows<double, int32_t>(a, *a.panel, pws, st)
                      : launch_panel_rows<double, int64_t>(a, *a.panel, pws, st))
               : (i32 ? launch_panel_rows<float, int32_t>(a, *a.panel, pws, st)
                      : launch_panel_rows<float, int64_t>(a, *a.panel, pws, st));
  };
  // Run one after the other: beside the unlisted rows' gathers from the
  // memory-side cache, the segments' B slices do not stay in L2 (measured:
  // the segment kernel 9.2 -> 17.7 ms, the SpMM 20.3 -> 22.5 ms).
  const hipError_t e = run_listed(s);
  return e != hipSuccess ? e : run_rows(s);
}

}  // namespace asc_hip
