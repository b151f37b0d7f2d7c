/*
 * Seeded synthetic CSR generators for the benchmark and test workloads.
 *
 * Every random draw is a pure function of (seed, stream, counter)
 * (SplitMix64 finalizer), and every row uses its own streams, so any row
 * range can be generated independently and the output does not depend on
 * the thread count.  Two phases: asc_gen_lengths() writes the int64 row
 * offsets and returns nnz; asc_gen_fill() writes sorted distinct column
 * indices (int32 or int64) and FP64 values.
 */
#ifndef ASC_SPARSE_GEN_H_
#define ASC_SPARSE_GEN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ASC_GEN_UNIFORM_FIXED = 0, /* per_row distinct uniform columns per row */
  ASC_GEN_ZIPF = 1,          /* row length Zipf(s) clipped to [1, zipf_max] */
  ASC_GEN_PRECISION = 2,     /* 16 within +-64 of the diagonal + 16 uniform */
  ASC_GEN_STENCIL2D = 3,     /* 5-point Laplacian on a k x k grid */
  ASC_GEN_STENCIL3D = 4,     /* 7-point Laplacian on a k x k x k grid */
  ASC_GEN_BANDED = 5,        /* half-bandwidth `band`, in-band kept w.p. fill */
  ASC_GEN_BLOCKDIAG8 = 6,    /* dense 8x8 diagonal blocks */
  ASC_GEN_BIPARTITE = 7,     /* rows link to a random half-window of columns */
  ASC_GEN_UNIFORM_VAR = 8    /* Poisson-like lengths around per_row, uniform */
};

enum {
  ASC_GEN_VALUES_UNIFORM = 0,   /* U[-1, 1] */
  ASC_GEN_VALUES_NORMAL = 1,    /* N(0, 1) */
  ASC_GEN_VALUES_LOGNORMAL = 2, /* random sign * exp(N(0, 2)) */
  ASC_GEN_VALUES_LAPLACIAN = 3  /* stencil weights: 2d (4, -1), 3d (6, -1) */
};

typedef struct asc_gen_params {
  int32_t family;
  int32_t value_dist;
  int64_t rows;
  int64_t cols;
  int64_t per_row;  /* UNIFORM_FIXED / UNIFORM_VAR / BIPARTITE degree */
  double zipf_s;
  int64_t zipf_max;
  int64_t band;     /* BANDED half bandwidth */
  double fill;      /* BANDED keep probability */
  int64_t grid;     /* STENCIL2D/3D grid edge k */
  uint64_t seed;
  int64_t window;   /* > 0: UNIFORM/ZIPF columns drawn from a window of this
                       width centred on the row's diagonal position */
} asc_gen_params;

/* Writes offsets[0..rows] and returns nnz (or -1 on a bad parameter). */
int64_t asc_gen_lengths(const asc_gen_params* p, int64_t* offsets, int threads);

/* Fills indices (index_bytes 4 or 8) and FP64 values for given offsets. */
int asc_gen_fill(const asc_gen_params* p, const int64_t* offsets, void* indices,
                 int index_bytes, double* values, int threads);

/* n draws of a value distribution (x vectors, dense blocks). */
void asc_gen_vector(int64_t n, uint64_t seed, int32_t value_dist, double* out,
                    int threads);

/* Sum of (rows+1) offsets etc. are the caller's; this helper returns the
 * expected mean of the clipped Zipf row length (for documentation/tests). */
double asc_gen_zipf_mean(double s, int64_t zipf_max);

#ifdef __cplusplus
}
#endif

#endif /* ASC_SPARSE_GEN_H_ */
