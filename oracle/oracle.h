/*
 * CPU oracle for the asc::sparse hot path (SpMV, SpMM, COO->CSR, CSR->CSC,
 * COO finalize).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C restatement of the
 * reference's serial Sparse BLAS and conversion algorithms.  It exists so the
 * test-suite and the benchmark's CPU-baseline leg have a checker; the product
 * (the HIP library under asc-cpp_amd/) never links, loads or calls it.
 *
 * Build rules (oracle/Makefile): gcc -O2 -ffp-contract=off, so every
 * `a + b * c` below is an unfused multiply then an add, as the reference's
 * serial code is when compiled by GCC for baseline x86-64.
 *
 * Index vocabulary follows the reference (include/asc/core/types.h:35-62):
 * every offset and index is signed 64-bit.
 *
 * Pinning (how each function is checked; see tests/test_oracle.py):
 *   oracle_compat_spmv_*    golden vectors of tests/sparse/blas_test.cc and
 *                           tests/sparse_cuda/sparse_cuda_test.cc, the
 *                           summation-order test, exact dyadic brute force,
 *                           and oracle/_ref (the reference compiled from its
 *                           own sources) on random inputs.
 *   oracle_std_spmv_*       golden vectors of standard_blas_test.cc, dense
 *                           brute force in exact integer arithmetic, and
 *                           oracle/_ref.
 *   oracle_std_spmm_*       as std_spmv, both layouts.
 *   oracle_*_faithful       the reference's O(outer*nnz) scan restated;
 *                           golden vectors of compressed_conversion_test.cc.
 *   oracle_*_fast           O(nnz) counting-sort restatement; equal to the
 *                           faithful form on random inputs (the canonical
 *                           order is unique) and to oracle/_ref.
 *   oracle_sparse_*         golden vectors of standard_blas_test.cc:69-174,
 *                           exact integer brute force, and oracle/_ref.
 *   oracle_trsv_*           golden vectors of standard_blas_test.cc:335-430
 *                           and oracle/_ref.
 *   oracle_coo_finalize_*   golden vectors of coordinate_test.cc and a
 *                           Python sorted()-based brute force.
 *   oracle_abs_spmv_*       (sets every plain-mode tolerance) exact integer
 *                           products, where the result must be exact; the
 *                           correctly rounded exact sum of |v||x| (Python
 *                           Fractions) within one unit in the last place on
 *                           random doubles; operands with cancelling signs,
 *                           where a dropped absolute value shows; and a
 *                           rectangular shape, where a transposed operand
 *                           shows.
 *   *_i32 variants          the same loop bodies over int32 indices; equal
 *                           bit for bit to the int64 form on random inputs.
 */
#ifndef ASC_SPARSE_HIP_ORACLE_H_
#define ASC_SPARSE_HIP_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status values reuse asc::ErrorCode numbering (include/asc/core/status.h:35-55). */
#define ORACLE_OK 0
#define ORACLE_INVALID_ARGUMENT 1
#define ORACLE_SHAPE 2
#define ORACLE_INDEX 3

/* Formats follow asc::SparseCompressedFormat (coordinate.h:86-89). */
#define ORACLE_CSR 0
#define ORACLE_CSC 1

/* Transposes follow asc::SparseBlasTranspose (blas.h:88-92). */
#define ORACLE_OP_N 0
#define ORACLE_OP_T 1
#define ORACLE_OP_H 2

/* Layouts follow asc::SparseBlasLayout (blas.h:103-106). */
#define ORACLE_COLUMN_MAJOR 0
#define ORACLE_ROW_MAJOR 1

/* Canonical structure check, restating compressed.h:711-746. */
int oracle_validate_compressed(int64_t outer, int64_t inner,
                               const int64_t* offsets, const int64_t* indices,
                               int64_t nnz);

/* Compat SpMV y = alpha*A*x (+ beta*y if beta != 0), reference_blas.cc:9-41. */
int oracle_compat_spmv_f64(int64_t rows, const int64_t* offsets,
                           const int64_t* indices, const double* values,
                           double alpha, const double* x, int64_t incx,
                           double beta, double* y, int64_t incy);
int oracle_compat_spmv_f32(int64_t rows, const int64_t* offsets,
                           const int64_t* indices, const float* values,
                           float alpha, const float* x, int64_t incx,
                           float beta, float* y, int64_t incy);
/* The same with int32 column indices (the device clones' narrow width). */
int oracle_compat_spmv_f64_i32(int64_t rows, const int64_t* offsets,
                               const int32_t* indices, const double* values,
                               double alpha, const double* x, int64_t incx,
                               double beta, double* y, int64_t incy);
int oracle_compat_spmv_f32_i32(int64_t rows, const int64_t* offsets,
                               const int32_t* indices, const float* values,
                               float alpha, const float* x, int64_t incx,
                               float beta, float* y, int64_t incy);

/* Standardized SpMV y += alpha*op(A)*x, blas.h:1607-1667 (loop :1241-1256). */
int oracle_std_spmv_f64(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const double* values, double alpha, const double* x,
                        int64_t incx, double* y, int64_t incy);
int oracle_std_spmv_f32(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const float* values, float alpha, const float* x,
                        int64_t incx, float* y, int64_t incy);

/* Standardized SpMM C += alpha*op(A)*B, blas.h:1692-1750. */
int oracle_std_spmm_f64(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const double* values, double alpha, const double* b,
                        int b_layout, int64_t ldb, int64_t k, double* c,
                        int c_layout, int64_t ldc);
int oracle_std_spmm_f32(int format, int transpose, int64_t rows, int64_t cols,
                        const int64_t* offsets, const int64_t* indices,
                        const float* values, float alpha, const float* b,
                        int b_layout, int64_t ldb, int64_t k, float* c,
                        int c_layout, int64_t ldc);
int oracle_std_spmm_f64_i32(int format, int transpose, int64_t rows, int64_t cols,
                            const int64_t* offsets, const int32_t* indices,
                            const double* values, double alpha, const double* b,
                            int b_layout, int64_t ldb, int64_t k, double* c,
                            int c_layout, int64_t ldc);

/* Per-row (|A||x|)_i in long double, for the plain-mode error bound
 * |y_gpu - y_ref| <= 4*len_i*eps*(|A||x|)_i.  Not a reference function. */
void oracle_abs_spmv_f64(int64_t rows, const int64_t* offsets,
                         const int64_t* indices, const double* values,
                         const double* x, double* out);
void oracle_abs_spmv_f64_i32(int64_t rows, const int64_t* offsets,
                             const int32_t* indices, const double* values,
                             const double* x, double* out);

/*
 * Canonical COO (interleaved int64 (row, col) pairs) to CSR (format 0) or
 * CSC (format 1).  Values are moved as raw bytes of width value_bytes.
 * faithful: compressed.h:1082-1131 (O(outer*nnz)).  fast: counting sort.
 */
int oracle_coo_to_compressed_faithful(int format, int64_t rows, int64_t cols,
                                      int64_t nnz, const int64_t* coords,
                                      const void* values, int value_bytes,
                                      int64_t* out_offsets,
                                      int64_t* out_indices, void* out_values);
int oracle_coo_to_compressed_fast(int format, int64_t rows, int64_t cols,
                                  int64_t nnz, const int64_t* coords,
                                  const void* values, int value_bytes,
                                  int64_t* out_offsets, int64_t* out_indices,
                                  void* out_values);

/*
 * CSR -> CSC (src_format 0) or CSC -> CSR (src_format 1).
 * faithful: compressed.h:1151-1221 (O(out_outer*(src_outer+nnz))).
 * fast: stable counting sort by inner index, O(nnz + outer).
 */
int oracle_transpose_faithful(int src_format, int64_t rows, int64_t cols,
                              const int64_t* offsets, const int64_t* indices,
                              const void* values, int value_bytes,
                              int64_t* out_offsets, int64_t* out_indices,
                              void* out_values);
int oracle_transpose_fast(int src_format, int64_t rows, int64_t cols,
                          const int64_t* offsets, const int64_t* indices,
                          const void* values, int value_bytes,
                          int64_t* out_offsets, int64_t* out_indices,
                          void* out_values);

/* Compressed -> canonical COO, compressed.h:1781-1837. */
int oracle_compressed_to_coo(int src_format, int64_t outer,
                             const int64_t* offsets, const int64_t* indices,
                             const void* values, int value_bytes,
                             int64_t* out_coords, void* out_values);

/*
 * CoordinateBuilder::Finalize for rank-2 double triplets,
 * coordinate.h:1309-1498: reject or sum duplicates (summed in insertion
 * order), then drop exact zeros after summation if asked (NaN is kept).
 * duplicate_policy: 0 reject, 1 sum.  zero_policy: 0 keep, 1 drop.
 * Returns ORACLE_INVALID_ARGUMENT (and writes nothing) when a duplicate is
 * rejected.  *out_nnz receives the finalized count.
 */
int oracle_coo_finalize_f64(int64_t nnz, const int64_t* coords,
                            const double* values, int duplicate_policy,
                            int zero_policy, int64_t* out_coords,
                            double* out_values, int64_t* out_nnz);
int oracle_coo_finalize_f32(int64_t nnz, const int64_t* coords,
                            const float* values, int duplicate_policy,
                            int zero_policy, int64_t* out_coords,
                            float* out_values, int64_t* out_nnz);

/* Sparse Level-1 over a canonical indexed vector (blas.h:1330-1582). */
double oracle_sparse_dot_f64(int64_t nnz, const int64_t* idx, const double* v,
                             const double* x, int64_t inc);
float oracle_sparse_dot_f32(int64_t nnz, const int64_t* idx, const float* v,
                            const float* x, int64_t inc);
void oracle_sparse_axpy_f64(int64_t nnz, double alpha, const int64_t* idx,
                            const double* v, double* x, int64_t inc);
void oracle_sparse_axpy_f32(int64_t nnz, float alpha, const int64_t* idx,
                            const float* v, float* x, int64_t inc);
void oracle_sparse_gather_f64(int64_t nnz, const int64_t* idx, double* v,
                              const double* x, int64_t inc);
void oracle_sparse_gather_f32(int64_t nnz, const int64_t* idx, float* v,
                              const float* x, int64_t inc);
void oracle_sparse_gather_zero_f64(int64_t nnz, const int64_t* idx, double* v,
                                   double* x, int64_t inc);
void oracle_sparse_gather_zero_f32(int64_t nnz, const int64_t* idx, float* v,
                                   float* x, int64_t inc);
void oracle_sparse_scatter_f64(int64_t nnz, const int64_t* idx, const double* v,
                               double* x, int64_t inc);
void oracle_sparse_scatter_f32(int64_t nnz, const int64_t* idx, const float* v,
                               float* x, int64_t inc);

/* SparseTriangularSolve(Multiple), blas.h:1760-1880, in place on k
 * right-hand sides (element (r, j) at b[r*rs + j*cs]).  lower_triangle: the
 * declared triangle is lower; unit: implicit unit diagonal. */
void oracle_trsv_f64(int fmt, int op, int lower_triangle, int unit, int64_t n,
                     const int64_t* off, const int64_t* idx, const double* val,
                     double alpha, double* b, int64_t rs, int64_t cs, int64_t k);
void oracle_trsv_f32(int fmt, int op, int lower_triangle, int unit, int64_t n,
                     const int64_t* off, const int64_t* idx, const float* val,
                     float alpha, float* b, int64_t rs, int64_t cs, int64_t k);

#ifdef __cplusplus
}
#endif

#endif /* ASC_SPARSE_HIP_ORACLE_H_ */
