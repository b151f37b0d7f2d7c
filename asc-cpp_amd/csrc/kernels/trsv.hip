// Sparse triangular solve for gfx950 (blas.h:1760-1880), level-scheduled.
//
// Analysis (once per matrix and operation): the stored matrix is checked
// against its declared triangle and, for a non-unit diagonal, for a stored
// nonzero diagonal (the reference's SparseBlasTriangularView::Create,
// blas.h:1090-1160).  Rows of op(A) are then levelled: a row's level is the
// round in which all rows it depends on were levelled in earlier rounds, one
// kernel launch per round, no spin-waiting.  A stable sort by level gives
// each level's rows in ascending order.
//
// Solve: one launch per level; one thread per (row, right-hand side) runs
// the row's serial chain  value = alpha*b;  value -= a*x  for each stored
// off-diagonal entry in column order;  value /= diagonal  (non-unit),
// without FMA.  This is the reference's dense loop restricted to stored
// entries: equal bit for bit while the solution stays finite (an unstored
// entry contributes an exact zero product there).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;

unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

// flags[0]: entry outside the declared triangle; flags[1]: missing or zero
// diagonal (non-unit only).
template <typename I, typename T>
__global__ __launch_bounds__(kBlock) void trsv_check_kernel(
    const int64_t* __restrict__ off, const I* __restrict__ idx, const T* __restrict__ val,
    int64_t outer, int inner_le_outer, int need_diag, uint32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t o = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; o < outer; o += stride) {
    bool diag = false;
    for (int64_t p = off[o]; p < off[o + 1]; ++p) {
      const int64_t in = static_cast<int64_t>(idx[p]);
      if (inner_le_outer ? in > o : in < o) flags[0] = 1u;
      if (in == o && val[p] != T(0)) diag = true;
    }
    if (need_diag && !diag) flags[1] = 1u;
  }
}

__global__ __launch_bounds__(kBlock) void trsv_level_init_kernel(int32_t* __restrict__ level,
                                                                 int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    level[i] = -1;
}

// One levelling round: a row not yet levelled whose dependencies were all
// levelled in earlier rounds gets level `round`.  flags[0]: some row was
// levelled; flags[1]: some row is still waiting.
template <typename I>
__global__ __launch_bounds__(kBlock) void trsv_level_round_kernel(
    const int64_t* __restrict__ off, const I* __restrict__ idx, int64_t n, int32_t round,
    int32_t* __restrict__ level, uint32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    if (level[i] >= 0) continue;
    bool ready = true;
    for (int64_t p = off[i]; p < off[i + 1] && ready; ++p) {
      const int64_t j = static_cast<int64_t>(idx[p]);
      if (j == i) continue;
      const int32_t lj = level[j];
      if (lj < 0 || lj >= round) ready = false;
    }
    if (ready) {
      level[i] = round;
      flags[0] = 1u;
    } else {
      flags[1] = 1u;
    }
  }
}

// One (row, right-hand side) of a level: the row's serial chain.
template <typename T, typename I>
__device__ __forceinline__ void solve_row(const int64_t* __restrict__ off,
                                          const I* __restrict__ idx, const T* __restrict__ val,
                                          int64_t row, int64_t j, int unit, T alpha,
                                          T* __restrict__ x, int64_t rs, int64_t cs) {
#pragma clang fp contract(off)
  T value = alpha * x[row * rs + j * cs];
  T diag = T(0);
  for (int64_t p = off[row]; p < off[row + 1]; ++p) {
    const int64_t col = static_cast<int64_t>(idx[p]);
    if (col == row) {
      diag = val[p];
      continue;
    }
    const T prod = val[p] * x[col * rs + j * cs];
    value = value - prod;
  }
  if (!unit) value = value / diag;
  x[row * rs + j * cs] = value;
}

template <typename T, typename I>
__global__ __launch_bounds__(kBlock) void trsv_level_solve_kernel(
    const int64_t* __restrict__ off, const I* __restrict__ idx, const T* __restrict__ val,
    const int64_t* __restrict__ level_ptr, const uint32_t* __restrict__ row_order, int64_t level,
    int unit, T alpha, T* __restrict__ x, int64_t rs, int64_t cs, int64_t k) {
  const int64_t begin = level_ptr[level];
  const int64_t count = level_ptr[level + 1] - begin;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= count * k) return;
  solve_row<T, I>(off, idx, val, row_order[begin + t % count], t / count, unit, alpha, x, rs, cs);
}

// Consecutive small levels [l0, l1) in one work-group: a barrier between
// levels orders each level's writes of x before the next level's reads (the
// work-group's waves share one CU and its L1), so a run of small levels
// costs one launch instead of one per level.
constexpr int kGroupBlock = 1024;
// Row solves of one level (times right-hand sides) that a grouped launch
// takes on: two passes of its threads.
constexpr int64_t kGroupMax = 2 * kGroupBlock;

template <typename T, typename I>
__global__ __launch_bounds__(kGroupBlock) void trsv_group_solve_kernel(
    const int64_t* __restrict__ off, const I* __restrict__ idx, const T* __restrict__ val,
    const int64_t* __restrict__ level_ptr, const uint32_t* __restrict__ row_order, int64_t l0,
    int64_t l1, int unit, T alpha, T* __restrict__ x, int64_t rs, int64_t cs, int64_t k) {
  for (int64_t l = l0; l < l1; ++l) {
    const int64_t begin = level_ptr[l];
    const int64_t count = level_ptr[l + 1] - begin;
    for (int64_t t = threadIdx.x; t < count * k; t += kGroupBlock)
      solve_row<T, I>(off, idx, val, row_order[begin + t % count], t / count, unit, alpha, x, rs,
                      cs);
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dense_zero_kernel(T* __restrict__ x, int64_t rows,
                                                            int64_t cols, int64_t rs, int64_t cs) {
  const int64_t n = rows * cols;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; t < n; t += stride)
    x[(t % rows) * rs + (t / rows) * cs] = T(0);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename I, typename T>
hipError_t check_typed(const Compressed& stored, bool inner_le_outer, bool unit, uint32_t* flags,
                       hipStream_t s) {
  trsv_check_kernel<I, T><<<grid_for(stored.outer), kBlock, 0, s>>>(
      stored.offsets, static_cast<const I*>(stored.indices), static_cast<const T*>(stored.values),
      stored.outer, inner_le_outer ? 1 : 0, unit ? 0 : 1, flags);
  return hipGetLastError();
}

// Levels of deep schedules in one work-group, in solve order: chunks of
// kDpBlock rows, each row's level from its dependencies outside the chunk
// (already final), then a Jacobi fixed point over the chunk in LDS:
//   level[i] = max(outside bound, 1 + level[j] for dependencies j in the chunk).
// Starting from the outside bounds the values only grow, never past the
// true levels, and a pass without change ends the chunk; a triangular
// chunk's longest internal chain has at most kDpBlock - 1 edges, so that
// happens within kDpBlock + 1 passes.  out[0] gets the number of levels.
constexpr int kDpBlock = 1024;

template <typename I>
__global__ __launch_bounds__(kDpBlock) void trsv_level_dp_kernel(
    const int64_t* __restrict__ off, const I* __restrict__ idx, int64_t n, int lower,
    int32_t* __restrict__ level, uint32_t* __restrict__ out) {
  __shared__ int32_t s_a[kDpBlock];
  __shared__ int32_t s_b[kDpBlock];
  __shared__ int32_t s_max[kDpBlock / kWave];
  __shared__ int s_changed;
  int32_t top = -1;
  const int64_t chunks = (n + kDpBlock - 1) / kDpBlock;
  for (int64_t c = 0; c < chunks; ++c) {
    const int64_t base = lower ? c * kDpBlock
                               : (n - (c + 1) * kDpBlock > 0 ? n - (c + 1) * kDpBlock : 0);
    const int64_t end = lower ? (base + kDpBlock < n ? base + kDpBlock : n) : n - c * kDpBlock;
    const int64_t i = base + threadIdx.x;
    const bool active = i < end;
    int32_t bound = 0;
    if (active)
      for (int64_t p = off[i]; p < off[i + 1]; ++p) {
        const int64_t j = static_cast<int64_t>(idx[p]);
        if (j == i || (j >= base && j < end)) continue;
        bound = level[j] + 1 > bound ? level[j] + 1 : bound;
      }
    s_a[threadIdx.x] = bound;
    int32_t* cur = s_a;
    int32_t* nxt = s_b;
    for (int pass = 0; pass <= kDpBlock; ++pass) {
      if (threadIdx.x == 0) s_changed = 0;
      __syncthreads();
      int32_t v = bound;
      if (active)
        for (int64_t p = off[i]; p < off[i + 1]; ++p) {
          const int64_t j = static_cast<int64_t>(idx[p]);
          if (j != i && j >= base && j < end) {
            const int32_t lj = cur[j - base] + 1;
            v = lj > v ? lj : v;
          }
        }
      nxt[threadIdx.x] = v;
      if (active && v != cur[threadIdx.x]) s_changed = 1;
      __syncthreads();
      const int changed = s_changed;
      int32_t* t = cur;
      cur = nxt;
      nxt = t;
      if (!changed) break;  // uniform: every thread read s_changed after the barrier
      __syncthreads();      // all reads of s_changed precede its next reset
    }
    if (active) {
      level[i] = cur[threadIdx.x];
      top = cur[threadIdx.x] > top ? cur[threadIdx.x] : top;
    }
    __syncthreads();  // this chunk's levels are visible to the next chunk's rows
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int32_t other = __shfl_xor(top, o);
    top = other > top ? other : top;
  }
  if (threadIdx.x % kWave == 0) s_max[threadIdx.x / kWave] = top;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t m = -1;
    for (int w = 0; w < kDpBlock / kWave; ++w) m = s_max[w] > m ? s_max[w] : m;
    out[0] = static_cast<uint32_t>(m + 1);
  }
}

// Shallow schedules level in a few synchronized multi-block rounds; a
// schedule still unfinished after kRounds rounds is levelled from scratch
// by the one-work-group pass above (no per-level host synchronization).
constexpr int32_t kRounds = 8;

template <typename I>
hipError_t level_typed(const Compressed& rows, bool lower, int32_t* level, uint32_t* flags,
                       uint32_t* host, int64_t* levels, hipStream_t s) {
  const int64_t n = rows.outer;
  *levels = 0;
  if (n == 0) return hipSuccess;
  trsv_level_init_kernel<<<grid_for(n), kBlock, 0, s>>>(level, n);
  hipError_t e = hipGetLastError();
  for (int32_t round = 0; e == hipSuccess && round < kRounds; ++round) {
    e = fill_launch(flags, 0, 8, s);
    if (e != hipSuccess) break;
    trsv_level_round_kernel<I><<<grid_for(n), kBlock, 0, s>>>(
        rows.offsets, static_cast<const I*>(rows.indices), n, round, level, flags);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipMemcpyAsync(host, flags, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
    if (host[1] == 0u) {  // every row levelled
      *levels = static_cast<int64_t>(round) + 1;
      return hipSuccess;
    }
    if (host[0] == 0u) return hipErrorInvalidValue;  // no progress: not triangular
  }
  if (e != hipSuccess) return e;
  trsv_level_dp_kernel<I><<<1, kDpBlock, 0, s>>>(rows.offsets, static_cast<const I*>(rows.indices),
                                                 n, lower ? 1 : 0, level, flags);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(host, flags, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  *levels = static_cast<int64_t>(host[0]);
  return hipSuccess;
}

}  // namespace

size_t trsv_analysis_scratch_bytes(int64_t n) {
  const size_t un = static_cast<size_t>(n > 0 ? n : 1);
  return 256 + align256(sizeof(int32_t) * un) + align256(sort_by_key_workspace_bytes(n, n + 1));
}

hipError_t trsv_analysis_launch(const Compressed& stored, bool stored_inner_le_outer, bool unit,
                                const Compressed& rows, bool lower, int64_t* level_ptr,
                                uint32_t* row_order, void* scratch, TrsvLevels* out,
                                int* failure, hipStream_t s) {
  *failure = 0;
  out->levels = 0;
  out->max_level_rows = 0;
  const int64_t n = rows.outer;
  char* ws = static_cast<char*>(scratch);
  uint32_t* flags = reinterpret_cast<uint32_t*>(ws);
  int32_t* level = reinterpret_cast<int32_t*>(ws + 256);
  char* sort_ws = ws + 256 + align256(sizeof(int32_t) * static_cast<size_t>(n > 0 ? n : 1));
  uint32_t host[2] = {0u, 0u};
  hipError_t e = fill_launch(flags, 0, 16, s);
  if (e != hipSuccess) return e;
  const bool i32 = stored.index_bytes == 4;
  if (stored.elem == Elem::F64)
    e = i32 ? check_typed<int32_t, double>(stored, stored_inner_le_outer, unit, flags, s)
            : check_typed<int64_t, double>(stored, stored_inner_le_outer, unit, flags, s);
  else
    e = i32 ? check_typed<int32_t, float>(stored, stored_inner_le_outer, unit, flags, s)
            : check_typed<int64_t, float>(stored, stored_inner_le_outer, unit, flags, s);
  if (e != hipSuccess) return e;
  if ((e = hipMemcpyAsync(host, flags, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (host[0]) {
    *failure = 1;
    return hipSuccess;
  }
  if (host[1]) {
    *failure = 2;
    return hipSuccess;
  }
  int64_t levels = 0;
  e = rows.index_bytes == 4 ? level_typed<int32_t>(rows, lower, level, flags, host, &levels, s)
                            : level_typed<int64_t>(rows, lower, level, flags, host, &levels, s);
  if (e != hipSuccess) return e;
  // Rows in level order (stable: ascending within a level); level_ptr gets
  // levels + 1 entries.
  e = sort_by_key_launch(reinterpret_cast<const uint32_t*>(level), n, levels, row_order, level_ptr,
                         sort_ws, s);
  if (e != hipSuccess) return e;
  // The level bounds come back to the host: the widest level sizes the
  // solve's grid, and the solve groups runs of small levels.
  std::vector<int64_t>& host_ptr = out->host_level_ptr;
  host_ptr.assign(static_cast<size_t>(levels + 1), 0);
  e = hipMemcpyAsync(host_ptr.data(), level_ptr, sizeof(int64_t) * (levels + 1),
                     hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  int64_t widest = 0;
  for (int64_t l = 0; e == hipSuccess && l < levels; ++l)
    widest = std::max(widest, host_ptr[l + 1] - host_ptr[l]);
  out->levels = levels;
  out->max_level_rows = widest;
  return e;
}

hipError_t trsv_solve_launch(const Compressed& rows, bool lower, bool unit, double alpha,
                             const int64_t* level_ptr, const uint32_t* row_order,
                             const TrsvLevels& lv, void* x, int64_t rs, int64_t cs, int64_t k,
                             hipStream_t s) {
  (void)lower;  // levels already order the solve
  if (lv.levels == 0 || k == 0) return hipSuccess;
  const int64_t threads = lv.max_level_rows * k;
  const unsigned grid = static_cast<unsigned>((threads + kBlock - 1) / kBlock);
  const bool i32 = rows.index_bytes == 4;
  const bool grouped = static_cast<int64_t>(lv.host_level_ptr.size()) == lv.levels + 1;
  for (int64_t l = 0; l < lv.levels; ++l) {
    if (grouped) {
      int64_t g = l;
      while (g < lv.levels && (lv.host_level_ptr[g + 1] - lv.host_level_ptr[g]) * k <= kGroupMax)
        ++g;
      if (g > l) {
#define ASC_TRSV_GROUP(T, I)                                                               \
  trsv_group_solve_kernel<T, I><<<1, kGroupBlock, 0, s>>>(                                 \
      rows.offsets, static_cast<const I*>(rows.indices), static_cast<const T*>(rows.values), \
      level_ptr, row_order, l, g, unit ? 1 : 0, static_cast<T>(alpha), static_cast<T*>(x),   \
      rs, cs, k)
        if (rows.elem == Elem::F64) {
          if (i32) ASC_TRSV_GROUP(double, int32_t);
          else ASC_TRSV_GROUP(double, int64_t);
        } else {
          if (i32) ASC_TRSV_GROUP(float, int32_t);
          else ASC_TRSV_GROUP(float, int64_t);
        }
#undef ASC_TRSV_GROUP
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        l = g - 1;
        continue;
      }
    }
    if (rows.elem == Elem::F64) {
      if (i32)
        trsv_level_solve_kernel<double, int32_t><<<grid, kBlock, 0, s>>>(
            rows.offsets, static_cast<const int32_t*>(rows.indices),
            static_cast<const double*>(rows.values), level_ptr, row_order, l, unit ? 1 : 0, alpha,
            static_cast<double*>(x), rs, cs, k);
      else
        trsv_level_solve_kernel<double, int64_t><<<grid, kBlock, 0, s>>>(
            rows.offsets, static_cast<const int64_t*>(rows.indices),
            static_cast<const double*>(rows.values), level_ptr, row_order, l, unit ? 1 : 0, alpha,
            static_cast<double*>(x), rs, cs, k);
    } else {
      if (i32)
        trsv_level_solve_kernel<float, int32_t><<<grid, kBlock, 0, s>>>(
            rows.offsets, static_cast<const int32_t*>(rows.indices),
            static_cast<const float*>(rows.values), level_ptr, row_order, l, unit ? 1 : 0,
            static_cast<float>(alpha), static_cast<float*>(x), rs, cs, k);
      else
        trsv_level_solve_kernel<float, int64_t><<<grid, kBlock, 0, s>>>(
            rows.offsets, static_cast<const int64_t*>(rows.indices),
            static_cast<const float*>(rows.values), level_ptr, row_order, l, unit ? 1 : 0,
            static_cast<float>(alpha), static_cast<float*>(x), rs, cs, k);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t dense_zero_launch(Elem elem, void* x, int64_t rows, int64_t cols, int64_t rs,
                             int64_t cs, hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (elem == Elem::F64)
    dense_zero_kernel<double><<<grid_for(rows * cols), kBlock, 0, s>>>(static_cast<double*>(x), rows,
                                                                       cols, rs, cs);
  else
    dense_zero_kernel<float><<<grid_for(rows * cols), kBlock, 0, s>>>(static_cast<float*>(x), rows,
                                                                      cols, rs, cs);
  return hipGetLastError();
}

}  // namespace asc_hip
