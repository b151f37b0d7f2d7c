/*
 * CPU oracle for the asc::sparse hot path.  TEST INFRASTRUCTURE ONLY: only
 * tests/, the smoke check and the benchmark's cpu_baseline leg load this
 * library, and only as a checker.  Compiled with -ffp-contract=off.
 *
 * Every function cites the reference file:line it restates; the header
 * (oracle.h) lists how each one is pinned.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Structure validation: include/asc/sparse/compressed.h:711-746.          */
/* offsets[0] == 0, nondecreasing, offsets[outer] == nnz, every inner index */
/* in [0, inner) and strictly increasing within its segment.                */
/* ------------------------------------------------------------------------ */
int oracle_validate_compressed(int64_t outer, int64_t inner,
                               const int64_t* offsets, const int64_t* indices,
                               int64_t nnz) {
  if (outer < 0 || inner < 0 || nnz < 0) return ORACLE_SHAPE;
  if (offsets[0] != 0 || offsets[outer] != nnz) return ORACLE_INVALID_ARGUMENT;
  for (int64_t o = 0; o < outer; ++o) {
    const int64_t b = offsets[o], e = offsets[o + 1];
    if (e < b || e > nnz) return ORACLE_INVALID_ARGUMENT;
    for (int64_t p = b; p < e; ++p) {
      if (indices[p] < 0 || indices[p] >= inner) return ORACLE_INDEX;
      if (p > b && indices[p] <= indices[p - 1]) return ORACLE_INVALID_ARGUMENT;
    }
  }
  return ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Compat SpMV: src/sparse/reference_blas.cc:9-41.                          */
/* product = +0; product += v*x in CSR order; result = alpha*product;       */
/* result += beta*y only when beta != 0 (y is not read otherwise).          */
/* ------------------------------------------------------------------------ */
#define COMPAT_SPMV_BODY(T)                                              \
  for (int64_t row = 0; row < rows; ++row) {                             \
    T product = (T)0;                                                    \
    for (int64_t p = offsets[row]; p < offsets[row + 1]; ++p) {          \
      product += values[p] * x[indices[p] * incx];                       \
    }                                                                    \
    T result = alpha * product;                                          \
    if (beta != (T)0) {                                                  \
      result += beta * y[row * incy];                                    \
    }                                                                    \
    y[row * incy] = result;                                              \
  }                                                                      \
  return ORACLE_OK;

int oracle_compat_spmv_f64(int64_t rows, const int64_t* offsets,
                           const int64_t* indices, const double* values,
                           double alpha, const double* x, int64_t incx,
                           double beta, double* y, int64_t incy) {
  COMPAT_SPMV_BODY(double)
}

int oracle_compat_spmv_f32(int64_t rows, const int64_t* offsets,
                           const int64_t* indices, const float* values,
                           float alpha, const float* x, int64_t incx,
                           float beta, float* y, int64_t incy) {
  COMPAT_SPMV_BODY(float)
}

/* The same loops over int32 column indices (the device clones' narrow
 * width), so large checks need no widened copy of the indices. */
int oracle_compat_spmv_f64_i32(int64_t rows, const int64_t* offsets,
                               const int32_t* indices, const double* values,
                               double alpha, const double* x, int64_t incx,
                               double beta, double* y, int64_t incy) {
  COMPAT_SPMV_BODY(double)
}

int oracle_compat_spmv_f32_i32(int64_t rows, const int64_t* offsets,
                               const int32_t* indices, const float* values,
                               float alpha, const float* x, int64_t incx,
                               float beta, float* y, int64_t incy) {
  COMPAT_SPMV_BODY(float)
}

/* ------------------------------------------------------------------------ */
/* Standardized SpMV: include/asc/sparse/blas.h:1607-1667, walking stored   */
/* entries with ForEachStored (:1241-1256): outer ascending, then position. */
/* For each (r, c, v): y[r*incy] += (alpha*v) * x[c*incx]; op != N swaps r  */
/* and c (conjugation is the identity for real scalars).  alpha == 0 is an  */
/* early return with no reads (:1650-1652).                                 */
/* ------------------------------------------------------------------------ */
#define STD_SPMV_BODY(T)                                                  \
  if (alpha == (T)0) return ORACLE_OK;                                    \
  const int64_t outer = format == ORACLE_CSR ? rows : cols;               \
  (void)cols;                                                             \
  for (int64_t o = 0; o < outer; ++o) {                                   \
    for (int64_t p = offsets[o]; p < offsets[o + 1]; ++p) {               \
      int64_t r = format == ORACLE_CSR ? o : indices[p];                  \
      int64_t c = format == ORACLE_CSR ? indices[p] : o;                  \
      if (transpose != ORACLE_OP_N) {                                     \
        const int64_t t = r;                                              \
        r = c;                                                            \
        c = t;                                                            \
      }                                                                   \
      y[r * incy] += alpha * values[p] * x[c * incx];                     \
    }                                                                     \
  }                                                                       \
  return ORACLE_OK;

int oracle_std_spmv_f64(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const double* values, double alpha, const double* x,
                        int64_t incx, double* y, int64_t incy) {
  STD_SPMV_BODY(double)
}

int oracle_std_spmv_f32(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const float* values, float alpha, const float* x,
                        int64_t incx, float* y, int64_t incy) {
  STD_SPMV_BODY(float)
}

/* ------------------------------------------------------------------------ */
/* Standardized SpMM: blas.h:1692-1750.  For each stored (r, c, v) in       */
/* canonical order and each rhs j: C(r,j) += (alpha*v) * B(c,j), with       */
/* element addressing of SparseBlasMatrixView::operator() (:1033-1038).     */
/* ------------------------------------------------------------------------ */
#define MAT_AT(ptr, layout, ld, r, c) \
  (ptr)[(layout) == ORACLE_COLUMN_MAJOR ? (c) * (ld) + (r) : (r) * (ld) + (c)]

#define STD_SPMM_BODY(T)                                                  \
  if (alpha == (T)0) return ORACLE_OK;                                    \
  const int64_t outer = format == ORACLE_CSR ? rows : cols;               \
  for (int64_t o = 0; o < outer; ++o) {                                   \
    for (int64_t p = offsets[o]; p < offsets[o + 1]; ++p) {               \
      int64_t r = format == ORACLE_CSR ? o : indices[p];                  \
      int64_t cc = format == ORACLE_CSR ? indices[p] : o;                 \
      if (transpose != ORACLE_OP_N) {                                     \
        const int64_t t = r;                                              \
        r = cc;                                                           \
        cc = t;                                                           \
      }                                                                   \
      const T v = values[p];                                              \
      for (int64_t j = 0; j < k; ++j) {                                   \
        MAT_AT(c, c_layout, ldc, r, j) +=                                 \
            alpha * v * MAT_AT(b, b_layout, ldb, cc, j);                  \
      }                                                                   \
    }                                                                     \
  }                                                                       \
  return ORACLE_OK;

int oracle_std_spmm_f64(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const double* values, double alpha, const double* b,
                        int b_layout, int64_t ldb, int64_t k, double* c,
                        int c_layout, int64_t ldc) {
  STD_SPMM_BODY(double)
}

int oracle_std_spmm_f32(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const float* values, float alpha, const float* b,
                        int b_layout, int64_t ldb, int64_t k, float* c,
                        int c_layout, int64_t ldc) {
  STD_SPMM_BODY(float)
}

int oracle_std_spmm_f64_i32(int format, int transpose, int64_t rows, int64_t cols,
                            const int64_t* offsets, const int32_t* indices,
                            const double* values, double alpha, const double* b,
                            int b_layout, int64_t ldb, int64_t k, double* c,
                            int c_layout, int64_t ldc) {
  STD_SPMM_BODY(double)
}

/* Row-wise (|A||x|)_i = sum_k |v_k| * |x_{c_k}|, accumulated in long double
 * (64-bit significand) and rounded once to double; test helper only. */
#define ABS_SPMV_BODY                                                        \
  for (int64_t r = 0; r < rows; ++r) {                                       \
    long double s = 0.0L;                                                    \
    for (int64_t p = offsets[r]; p < offsets[r + 1]; ++p) {                  \
      s += fabsl((long double)values[p]) * fabsl((long double)x[indices[p]]); \
    }                                                                        \
    out[r] = (double)s;                                                      \
  }

void oracle_abs_spmv_f64(int64_t rows, const int64_t* offsets,
                         const int64_t* indices, const double* values,
                         const double* x, double* out) {
  ABS_SPMV_BODY
}

void oracle_abs_spmv_f64_i32(int64_t rows, const int64_t* offsets,
                             const int32_t* indices, const double* values,
                             const double* x, double* out) {
  ABS_SPMV_BODY
}

/* ------------------------------------------------------------------------ */
/* COO -> compressed.                                                       */
/* ------------------------------------------------------------------------ */

/* compressed.h:1082-1131: for each output outer index, scan every stored   */
/* coordinate in COO order and append those whose outer coordinate matches. */
int oracle_coo_to_compressed_faithful(int format, int64_t rows, int64_t cols,
                                      int64_t nnz, const int64_t* coords,
                                      const void* values, int value_bytes,
                                      int64_t* out_offsets,
                                      int64_t* out_indices, void* out_values) {
  const int64_t outer_extent = format == ORACLE_CSR ? rows : cols;
  const char* src = (const char*)values;
  char* dst = (char*)out_values;
  int64_t output = 0;
  out_offsets[0] = 0;
  for (int64_t outer = 0; outer < outer_extent; ++outer) {
    for (int64_t p = 0; p < nnz; ++p) {
      const int64_t r = coords[2 * p], c = coords[2 * p + 1];
      const int64_t candidate = format == ORACLE_CSR ? r : c;
      if (candidate != outer) continue;
      out_indices[output] = format == ORACLE_CSR ? c : r;
      memcpy(dst + output * value_bytes, src + p * value_bytes,
             (size_t)value_bytes);
      ++output;
    }
    out_offsets[outer + 1] = output;
  }
  return ORACLE_OK;
}

/* The same output by a stable counting sort on the outer coordinate: count, */
/* exclusive scan, then place entries in COO order.  Stability reproduces   */
/* the faithful scan's within-bucket order (COO order).                     */
int oracle_coo_to_compressed_fast(int format, int64_t rows, int64_t cols,
                                  int64_t nnz, const int64_t* coords,
                                  const void* values, int value_bytes,
                                  int64_t* out_offsets, int64_t* out_indices,
                                  void* out_values) {
  const int64_t outer_extent = format == ORACLE_CSR ? rows : cols;
  const int oi = format == ORACLE_CSR ? 0 : 1;
  const char* src = (const char*)values;
  char* dst = (char*)out_values;
  memset(out_offsets, 0, (size_t)(outer_extent + 1) * sizeof(int64_t));
  for (int64_t p = 0; p < nnz; ++p) out_offsets[coords[2 * p + oi] + 1]++;
  for (int64_t o = 0; o < outer_extent; ++o) out_offsets[o + 1] += out_offsets[o];
  int64_t* cursor = (int64_t*)malloc((size_t)(outer_extent + 1) * sizeof(int64_t));
  if (!cursor) return ORACLE_SHAPE;
  memcpy(cursor, out_offsets, (size_t)(outer_extent + 1) * sizeof(int64_t));
  for (int64_t p = 0; p < nnz; ++p) {
    const int64_t o = coords[2 * p + oi];
    const int64_t d = cursor[o]++;
    out_indices[d] = coords[2 * p + (1 - oi)];
    memcpy(dst + d * value_bytes, src + p * value_bytes, (size_t)value_bytes);
  }
  free(cursor);
  return ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Compressed transpose (CSR <-> CSC).                                      */
/* ------------------------------------------------------------------------ */

/* compressed.h:1151-1221: for each output outer index, scan the source in  */
/* source-outer then position order and append matching entries.           */
int oracle_transpose_faithful(int src_format, int64_t rows, int64_t cols,
                              const int64_t* offsets, const int64_t* indices,
                              const void* values, int value_bytes,
                              int64_t* out_offsets, int64_t* out_indices,
                              void* out_values) {
  const int64_t src_outer = src_format == ORACLE_CSR ? rows : cols;
  const int64_t dst_outer = src_format == ORACLE_CSR ? cols : rows;
  const char* src = (const char*)values;
  char* dst = (char*)out_values;
  int64_t output = 0;
  out_offsets[0] = 0;
  for (int64_t oo = 0; oo < dst_outer; ++oo) {
    for (int64_t so = 0; so < src_outer; ++so) {
      for (int64_t p = offsets[so]; p < offsets[so + 1]; ++p) {
        if (indices[p] != oo) continue;
        out_indices[output] = so;
        memcpy(dst + output * value_bytes, src + p * value_bytes,
               (size_t)value_bytes);
        ++output;
      }
    }
    out_offsets[oo + 1] = output;
  }
  return ORACLE_OK;
}

/* Stable counting sort by inner index over the source's canonical order.  */
int oracle_transpose_fast(int src_format, int64_t rows, int64_t cols,
                          const int64_t* offsets, const int64_t* indices,
                          const void* values, int value_bytes,
                          int64_t* out_offsets, int64_t* out_indices,
                          void* out_values) {
  const int64_t src_outer = src_format == ORACLE_CSR ? rows : cols;
  const int64_t dst_outer = src_format == ORACLE_CSR ? cols : rows;
  const int64_t nnz = offsets[src_outer];
  const char* src = (const char*)values;
  char* dst = (char*)out_values;
  memset(out_offsets, 0, (size_t)(dst_outer + 1) * sizeof(int64_t));
  for (int64_t p = 0; p < nnz; ++p) out_offsets[indices[p] + 1]++;
  for (int64_t o = 0; o < dst_outer; ++o) out_offsets[o + 1] += out_offsets[o];
  int64_t* cursor = (int64_t*)malloc((size_t)(dst_outer + 1) * sizeof(int64_t));
  if (!cursor) return ORACLE_SHAPE;
  memcpy(cursor, out_offsets, (size_t)(dst_outer + 1) * sizeof(int64_t));
  for (int64_t so = 0; so < src_outer; ++so) {
    for (int64_t p = offsets[so]; p < offsets[so + 1]; ++p) {
      const int64_t d = cursor[indices[p]]++;
      out_indices[d] = so;
      memcpy(dst + d * value_bytes, src + p * value_bytes, (size_t)value_bytes);
    }
  }
  free(cursor);
  return ORACLE_OK;
}

/* compressed.h:1781-1837: emit (row, col) for each stored entry in outer  */
/* then position order.  For CSC input the result is re-sorted row-major  */
/* (canonical COO is lexicographic by (row, col)), i.e. a CSC->CSR         */
/* transpose followed by the CSR walk.                                     */
int oracle_compressed_to_coo(int src_format, int64_t outer,
                             const int64_t* offsets, const int64_t* indices,
                             const void* values, int value_bytes,
                             int64_t* out_coords, void* out_values) {
  const char* src = (const char*)values;
  char* dst = (char*)out_values;
  if (src_format == ORACLE_CSR) {
    for (int64_t r = 0; r < outer; ++r) {
      for (int64_t p = offsets[r]; p < offsets[r + 1]; ++p) {
        out_coords[2 * p] = r;
        out_coords[2 * p + 1] = indices[p];
        memcpy(dst + p * value_bytes, src + p * value_bytes,
               (size_t)value_bytes);
      }
    }
    return ORACLE_OK;
  }
  /* CSC: rows are the inner index; find the row extent. */
  const int64_t nnz = offsets[outer];
  int64_t rows = 0;
  for (int64_t p = 0; p < nnz; ++p)
    if (indices[p] + 1 > rows) rows = indices[p] + 1;
  int64_t* off = (int64_t*)malloc((size_t)(rows + 1) * sizeof(int64_t));
  int64_t* idx = (int64_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
  char* val = (char*)malloc((size_t)(nnz > 0 ? nnz : 1) * (size_t)value_bytes);
  if (!off || !idx || !val) {
    free(off);
    free(idx);
    free(val);
    return ORACLE_SHAPE;
  }
  oracle_transpose_fast(ORACLE_CSC, rows, outer, offsets, indices, values,
                        value_bytes, off, idx, val);
  const int rc = oracle_compressed_to_coo(ORACLE_CSR, rows, off, idx, val,
                                          value_bytes, out_coords, out_values);
  free(off);
  free(idx);
  free(val);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* CoordinateBuilder::Finalize, coordinate.h:1309-1498.                     */
/* PreflightDuplicates (:1391-1428) rejects before any change; StableSort   */
/* (:1430-1461) is a stable sort by (row, col); Compact (:1463-1498) sums  */
/* equal coordinates in their stable order (= insertion order) and applies */
/* the zero policy after summation (NaN != 0, so NaN is kept).  A stable   */
/* merge sort gives the same permutation as the stable insertion sort.     */
/* ------------------------------------------------------------------------ */
static const int64_t* g_sort_coords;

static int coord_less(int64_t a, int64_t b) {
  const int64_t ra = g_sort_coords[2 * a], rb = g_sort_coords[2 * b];
  if (ra != rb) return ra < rb;
  return g_sort_coords[2 * a + 1] < g_sort_coords[2 * b + 1];
}

static void merge_sort(int64_t* perm, int64_t* tmp, int64_t n) {
  for (int64_t width = 1; width < n; width *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * width) {
      int64_t mid = lo + width < n ? lo + width : n;
      int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
      int64_t i = lo, j = mid, o = lo;
      while (i < mid && j < hi) {
        /* take from the right only if strictly less: stability */
        if (coord_less(perm[j], perm[i])) tmp[o++] = perm[j++];
        else tmp[o++] = perm[i++];
      }
      while (i < mid) tmp[o++] = perm[i++];
      while (j < hi) tmp[o++] = perm[j++];
    }
    memcpy(perm, tmp, (size_t)n * sizeof(int64_t));
  }
}

/* One body for both element types: T is the element, NAME the symbol. */
#define ORACLE_FINALIZE_BODY(NAME, T)                                            \
  int NAME(int64_t nnz, const int64_t* coords, const T* values,                 \
           int duplicate_policy, int zero_policy, int64_t* out_coords,          \
           T* out_values, int64_t* out_nnz) {                                   \
    int64_t* perm = (int64_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t)); \
    int64_t* tmp = (int64_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));  \
    if (!perm || !tmp) {                                                         \
      free(perm);                                                                \
      free(tmp);                                                                 \
      return ORACLE_SHAPE;                                                       \
    }                                                                            \
    for (int64_t i = 0; i < nnz; ++i) perm[i] = i;                               \
    g_sort_coords = coords;                                                      \
    merge_sort(perm, tmp, nnz);                                                  \
    /* Preflight: any duplicate under kReject fails before output is written. */ \
    if (duplicate_policy == 0) {                                                 \
      for (int64_t i = 1; i < nnz; ++i) {                                        \
        const int64_t a = perm[i - 1], b = perm[i];                              \
        if (coords[2 * a] == coords[2 * b] && coords[2 * a + 1] == coords[2 * b + 1]) { \
          free(perm);                                                            \
          free(tmp);                                                             \
          *out_nnz = 0;                                                          \
          return ORACLE_INVALID_ARGUMENT;                                        \
        }                                                                        \
      }                                                                          \
    }                                                                            \
    int64_t output = 0, input = 0;                                               \
    while (input < nnz) {                                                        \
      const int64_t head = perm[input++];                                        \
      T value = values[head];                                                    \
      while (input < nnz && coords[2 * perm[input]] == coords[2 * head] &&       \
             coords[2 * perm[input] + 1] == coords[2 * head + 1]) {              \
        value += values[perm[input]];                                            \
        ++input;                                                                 \
      }                                                                          \
      if (zero_policy == 1 && value == (T)0) continue;                           \
      out_coords[2 * output] = coords[2 * head];                                 \
      out_coords[2 * output + 1] = coords[2 * head + 1];                         \
      out_values[output] = value;                                                \
      ++output;                                                                  \
    }                                                                            \
    *out_nnz = output;                                                           \
    free(perm);                                                                  \
    free(tmp);                                                                   \
    return ORACLE_OK;                                                            \
  }

ORACLE_FINALIZE_BODY(oracle_coo_finalize_f64, double)
ORACLE_FINALIZE_BODY(oracle_coo_finalize_f32, float)

/* ------------------------------------------------------------------------ */
/* Sparse Level-1 (blas.h:1330-1582) over a canonical indexed vector        */
/* (strictly increasing indices, blas.h:606-620) and a strided dense vector */
/* (element i at dense[i*inc], inc > 0 here).  Real elements: conjugation  */
/* is the identity.                                                         */
/*   dot:         r = +0; r += v[p]*x[idx[p]] for p in order (:1360-1366)   */
/*   axpy:        alpha == 0 returns; x[idx[p]] += alpha*v[p] (:1391-1398)  */
/*   gather:      v[p] = x[idx[p]] (:1427-1430)                             */
/*   gather-zero: v[p] = x[idx[p]]; x[idx[p]] = 0 (:1460-1465)              */
/*   scatter:     x[idx[p]] = v[p] (:1494-1497)                             */
/* ------------------------------------------------------------------------ */
#define ORACLE_LEVEL1(SUFFIX, T)                                                 \
  T oracle_sparse_dot_##SUFFIX(int64_t nnz, const int64_t* idx, const T* v,     \
                               const T* x, int64_t inc) {                       \
    T r = (T)0;                                                                  \
    for (int64_t p = 0; p < nnz; ++p) r += v[p] * x[idx[p] * inc];               \
    return r;                                                                    \
  }                                                                              \
  void oracle_sparse_axpy_##SUFFIX(int64_t nnz, T alpha, const int64_t* idx,    \
                                   const T* v, T* x, int64_t inc) {             \
    if (alpha == (T)0) return;                                                   \
    for (int64_t p = 0; p < nnz; ++p) x[idx[p] * inc] += alpha * v[p];           \
  }                                                                              \
  void oracle_sparse_gather_##SUFFIX(int64_t nnz, const int64_t* idx, T* v,     \
                                     const T* x, int64_t inc) {                 \
    for (int64_t p = 0; p < nnz; ++p) v[p] = x[idx[p] * inc];                    \
  }                                                                              \
  void oracle_sparse_gather_zero_##SUFFIX(int64_t nnz, const int64_t* idx,      \
                                          T* v, T* x, int64_t inc) {            \
    for (int64_t p = 0; p < nnz; ++p) {                                          \
      v[p] = x[idx[p] * inc];                                                    \
      x[idx[p] * inc] = (T)0;                                                    \
    }                                                                            \
  }                                                                              \
  void oracle_sparse_scatter_##SUFFIX(int64_t nnz, const int64_t* idx,          \
                                      const T* v, T* x, int64_t inc) {          \
    for (int64_t p = 0; p < nnz; ++p) x[idx[p] * inc] = v[p];                    \
  }

ORACLE_LEVEL1(f64, double)
ORACLE_LEVEL1(f32, float)

/* ------------------------------------------------------------------------ */
/* SparseTriangularSolve(Multiple), blas.h:1760-1880: in place on k          */
/* right-hand sides (element (r, j) at b[r*rs + j*cs]).  alpha == 0 zeroes  */
/* the right-hand sides.  lower = transposed ? triangle == upper :          */
/* triangle == lower; rows run forward (lower) or backward (upper); each    */
/* row subtracts op(A)(row, col)*x(col) for every col strictly before it in */
/* the solve order, in increasing col order, as the reference's dense loop  */
/* does (unstored entries contribute 0*x), then divides by the diagonal     */
/* when non-unit.  op(A)(r, c) is looked up in the stored CSR/CSC arrays.   */
/* ------------------------------------------------------------------------ */
#define ORACLE_TRSV(SUFFIX, T)                                                   \
  static T op_value_##SUFFIX(int fmt, int op, const int64_t* off,               \
                             const int64_t* idx, const T* val, int64_t r,       \
                             int64_t c) {                                       \
    /* op(A)(r, c) = A(r, c) for N, A(c, r) for T/H (real: no conjugation) */   \
    const int64_t ar = op == ORACLE_OP_N ? r : c;                                \
    const int64_t ac = op == ORACLE_OP_N ? c : r;                                \
    const int64_t outer = fmt == ORACLE_CSR ? ar : ac;                           \
    const int64_t inner = fmt == ORACLE_CSR ? ac : ar;                           \
    for (int64_t p = off[outer]; p < off[outer + 1]; ++p)                        \
      if (idx[p] == inner) return val[p];                                        \
    return (T)0;                                                                 \
  }                                                                              \
  void oracle_trsv_##SUFFIX(int fmt, int op, int lower_triangle, int unit,      \
                            int64_t n, const int64_t* off, const int64_t* idx,  \
                            const T* val, T alpha, T* b, int64_t rs, int64_t cs, \
                            int64_t k) {                                        \
    if (alpha == (T)0) {                                                         \
      for (int64_t j = 0; j < k; ++j)                                            \
        for (int64_t r = 0; r < n; ++r) b[r * rs + j * cs] = (T)0;               \
      return;                                                                    \
    }                                                                            \
    const int transposed = op != ORACLE_OP_N;                                    \
    const int lower = transposed ? !lower_triangle : lower_triangle;             \
    for (int64_t j = 0; j < k; ++j) {                                            \
      for (int64_t step = 0; step < n; ++step) {                                 \
        const int64_t row = lower ? step : n - 1 - step;                         \
        T value = alpha * b[row * rs + j * cs];                                  \
        for (int64_t col = 0; col < n; ++col) {                                  \
          if ((lower && col >= row) || (!lower && col <= row)) continue;         \
          const T prod = op_value_##SUFFIX(fmt, op, off, idx, val, row, col) *   \
                         b[col * rs + j * cs];                                   \
          value -= prod;                                                         \
        }                                                                        \
        if (!unit) value /= op_value_##SUFFIX(fmt, op, off, idx, val, row, row); \
        b[row * rs + j * cs] = value;                                            \
      }                                                                          \
    }                                                                            \
  }

ORACLE_TRSV(f64, double)
ORACLE_TRSV(f32, float)
