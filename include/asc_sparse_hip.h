/*
 * asc_sparse_hip.h -- C ABI of ASC::sparse_hip, an MI355X (gfx950) Sparse
 * BLAS provider for the asc::sparse CSR/CSC/COO storage types.
 *
 * This is the type-erased seam that the C++ facet
 * (include/asc/sparse/providers/hip.h) and foreign callers (ctypes, cgo, JNI)
 * bind.  It replaces the reference's exported type-erased provider functions
 * (include/asc/sparse/providers/cuda.h:212-253) one for one where the
 * reference has a counterpart, and adds the device conversions the reference
 * provider lacks.
 *
 * Conventions
 *  - Every function returns asc_hip_status.  `code` carries the numeric value
 *    of asc::ErrorCode (include/asc/core/status.h:35-55); `native` carries the
 *    hipError_t when `provider` is ASC_HIP_PROVIDER_HIP.  A human-readable
 *    message for the calling thread's last failure is available from
 *    asc_sparse_hip_last_message().  Nothing throws across this boundary.
 *  - Every operand is validated before anything is enqueued; a failing call
 *    leaves every output untouched (reference operations.cc:349-398).
 *  - Operations are stream-ordered on the context's HIP stream and return
 *    without waiting; record an event or synchronize the context to observe
 *    completion (reference runtime.cc:467-482).
 *  - Operations never allocate device memory.  Temporaries come from a
 *    caller-owned workspace whose size is queried first (reference
 *    providers/cuda.h:1758-1771, docs/modules/sparse.md:481-486).
 *  - Device matrices must be "trusted": produced by asc_sparse_hip_clone_* or
 *    a device conversion of this library (reference compressed.h:759-770,
 *    operations.cc:163-166).
 *  - One context must not be used from two threads at once; independent
 *    contexts may run concurrently (reference docs/modules/sparse.md:440-442).
 */
#ifndef ASC_SPARSE_HIP_H_
#define ASC_SPARSE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */

/* asc::ErrorCode values (include/asc/core/status.h:35-55). */
enum {
  ASC_HIP_OK = 0,
  ASC_HIP_INVALID_ARGUMENT = 1,
  ASC_HIP_SHAPE = 2,
  ASC_HIP_INDEX = 3,
  ASC_HIP_OVERFLOW = 4,
  ASC_HIP_INVALID_STATE = 5,
  ASC_HIP_ALLOCATION = 6,
  ASC_HIP_MEMORY_ACCESS = 7,
  ASC_HIP_MEMORY_TRANSFER = 8,
  ASC_HIP_UNSUPPORTED = 9,
  ASC_HIP_UNAVAILABLE = 10,
  ASC_HIP_PROVIDER = 11,
  ASC_HIP_NUMERICAL = 12,
  ASC_HIP_INTERNAL = 18
};

enum { ASC_HIP_PROVIDER_NONE = 0, ASC_HIP_PROVIDER_HIP = 1 };

typedef struct asc_hip_status {
  int32_t code;     /* asc::ErrorCode */
  int32_t provider; /* ASC_HIP_PROVIDER_* */
  int64_t native;   /* hipError_t when provider == ASC_HIP_PROVIDER_HIP */
} asc_hip_status;

/* ---- vocabulary (values mirror the reference enums) -------------------- */

/* internal_sparse_cuda::ElementKind (providers/cuda.h:103-108), extended by
 * the storage-only reduced precisions of this provider. */
enum {
  ASC_HIP_F32 = 0,
  ASC_HIP_F64 = 1,
  ASC_HIP_C64 = 2,  /* reserved: complex float */
  ASC_HIP_C128 = 3, /* reserved: complex double */
  ASC_HIP_BF16 = 4, /* storage only; arithmetic in FP32 */
  ASC_HIP_F16 = 5   /* storage only; arithmetic in FP32 */
};

/* internal_sparse_cuda::FormatKind (providers/cuda.h:125-129). */
enum { ASC_HIP_COO = 0, ASC_HIP_CSR = 1, ASC_HIP_CSC = 2 };

/* asc::MemorySpace (include/asc/core/memory.h:30-35). */
enum {
  ASC_HIP_HOST = 0,
  ASC_HIP_PINNED_HOST = 1,
  ASC_HIP_DEVICE = 2,
  ASC_HIP_MANAGED = 3
};

/* asc::SparseBlasTranspose (include/asc/sparse/blas.h:88-92). */
enum { ASC_HIP_OP_N = 0, ASC_HIP_OP_T = 1, ASC_HIP_OP_H = 2 };

/* asc::SparseBlasLayout (include/asc/sparse/blas.h:103-106). */
enum { ASC_HIP_COLUMN_MAJOR = 0, ASC_HIP_ROW_MAJOR = 1 };

/* asc::Determinism (include/asc/core/execution.h:122-125). */
enum { ASC_HIP_DETERMINISTIC = 0, ASC_HIP_BACKEND_DEFAULT = 1 };

/* asc::DuplicatePolicy / ExplicitZeroPolicy (coordinate.h:60-78). */
enum { ASC_HIP_DUPLICATE_REJECT = 0, ASC_HIP_DUPLICATE_SUM = 1 };
enum { ASC_HIP_ZERO_KEEP = 0, ASC_HIP_ZERO_DROP = 1 };

/*
 * Sparse operand, after internal_sparse_cuda::SparseDescriptor
 * (providers/cuda.h:143-156).  CSR: outer = rows; CSC: outer = cols.
 * COO: `outer_offsets` points at the interleaved (row, col) int64
 * coordinates and `inner_indices` is unused.  Offsets are int64
 * (offset_bytes = 8; 4 is accepted for offsets that fit int32); inner indices
 * are int64 as in the reference or int32 after an explicit narrowing clone.
 */
typedef struct asc_hip_sparse_desc {
  const void* outer_offsets;
  const void* inner_indices;
  const void* values;
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  uint8_t format;       /* ASC_HIP_COO / CSR / CSC */
  uint8_t element_kind; /* ASC_HIP_F32 ... */
  uint8_t index_bytes;  /* 4 or 8 */
  uint8_t offset_bytes; /* 4 or 8 */
  uint8_t memory_space; /* ASC_HIP_HOST ... */
  uint8_t trusted;      /* canonical structure vouched for by this library */
  uint8_t reserved[2];
} asc_hip_sparse_desc;

/* Strided vector, after VectorDescriptor (providers/cuda.h:158-167).
 * `data` addresses logical element 0; element i is data[i * stride]. */
typedef struct asc_hip_vector_desc {
  const void* data;
  int64_t extent;
  int64_t stride;
  uint8_t element_kind;
  uint8_t memory_space;
  uint8_t writable;
  uint8_t reserved[5];
} asc_hip_vector_desc;

/* Dense block, after MatrixDescriptor (providers/cuda.h:180-191). */
typedef struct asc_hip_matrix_desc {
  const void* data;
  int64_t rows;
  int64_t cols;
  int64_t leading_dimension;
  uint8_t layout; /* ASC_HIP_COLUMN_MAJOR / ROW_MAJOR */
  uint8_t element_kind;
  uint8_t memory_space;
  uint8_t writable;
  uint8_t reserved[4];
} asc_hip_matrix_desc;

/* ScalarValue (providers/cuda.h:119-122). */
typedef struct asc_hip_scalar {
  double real;
  double imaginary;
} asc_hip_scalar;

/*
 * SpMV/SpMM algorithm choice.  ASC_HIP_ALG_AUTO uses a tuned or learned
 * choice when the context holds one for the matrix and the merge-path kernel
 * otherwise.  ASC_HIP_ALG_ORDERED reproduces the serial reference's
 * left-to-right per-row summation bit for bit (no FMA, no reassociation).
 * The remaining values name one kernel variant each (the autotuner's space).
 */
enum {
  ASC_HIP_ALG_AUTO = 0,
  ASC_HIP_ALG_ORDERED = 1,
  ASC_HIP_ALG_MERGE = 2,    /* merge-path: rows+nnz split evenly per tile */
  ASC_HIP_ALG_VECTOR1 = 3,  /* 1 lane per row */
  ASC_HIP_ALG_VECTOR2 = 4,  /* 2 lanes per row */
  ASC_HIP_ALG_VECTOR4 = 5,
  ASC_HIP_ALG_VECTOR8 = 6,
  ASC_HIP_ALG_VECTOR16 = 7,
  ASC_HIP_ALG_VECTOR32 = 8,
  ASC_HIP_ALG_VECTOR64 = 9, /* one wave64 per row */
  ASC_HIP_ALG_PANEL = 10,   /* column panels; needs asc_sparse_hip_csr_spmv_analyze */
  ASC_HIP_ALG_COUNT = 11
};

/* Accumulation for FP32/BF16/F16-stored values (FP64 always accumulates in
 * FP64).  TWOSUM is compensated (Knuth TwoSum plus the FMA product error). */
enum { ASC_HIP_ACC_PLAIN = 0, ASC_HIP_ACC_TWOSUM = 1 };

typedef struct asc_hip_spmv_options {
  uint8_t algorithm;  /* ASC_HIP_ALG_* */
  uint8_t accumulate; /* ASC_HIP_ACC_* */
  uint8_t reserved[6];
} asc_hip_spmv_options;

typedef struct asc_hip_context asc_hip_context;
typedef struct asc_hip_event asc_hip_event;

/* ---- context, memory, events (core_hip) -------------------------------- */

/* Binds `device` and a stream.  `stream` is a hipStream_t to borrow, or NULL
 * to create a non-blocking stream owned by the context.  Mirrors
 * SparseCudaContext::Create (providers/cuda.h:514; context.cc:69-113) over
 * CreateCudaExecutionContext (src/core/cuda/runtime.cc:434-465). */
asc_hip_status asc_sparse_hip_context_create(int32_t device, void* stream,
                                             uint8_t determinism,
                                             asc_hip_context** out);
asc_hip_status asc_sparse_hip_context_destroy(asc_hip_context* context);
asc_hip_status asc_sparse_hip_context_stream(asc_hip_context* context,
                                             void** out_stream);
/* Waits for the context's stream.  Also the point where a kernel that gave
 * up a bounded wait is reported (once): ASC_HIP_INTERNAL, and the outputs of
 * the operations since the previous synchronization are not valid. */
asc_hip_status asc_sparse_hip_context_synchronize(asc_hip_context* context);

/* Message of the calling thread's most recent failed call ("" if none). */
const char* asc_sparse_hip_last_message(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int32_t asc_sparse_hip_device_count(void);

/* Device / pinned-host allocation, after CudaMemoryResource
 * (include/asc/core/providers/cuda.h:39-120).  Zero bytes yields NULL. */
asc_hip_status asc_sparse_hip_allocate(asc_hip_context* context, size_t bytes,
                                       uint8_t memory_space, void** out);
asc_hip_status asc_sparse_hip_deallocate(asc_hip_context* context, void* ptr,
                                         uint8_t memory_space);
/* Stream-ordered copy between any host/device pointers (CopyBytes,
 * src/core/cuda/runtime.cc:226-253). */
asc_hip_status asc_sparse_hip_copy(asc_hip_context* context, void* dst,
                                   const void* src, size_t bytes);
/* Stream-ordered byte fill (a kernel, safe under graph capture). */
asc_hip_status asc_sparse_hip_fill(asc_hip_context* context, void* dst,
                                   uint8_t byte, size_t bytes);

/* Completion events (CompletionEvent, include/asc/core/execution.h:246-369). */
asc_hip_status asc_sparse_hip_event_record(asc_hip_context* context,
                                           asc_hip_event** out);
asc_hip_status asc_sparse_hip_event_wait(asc_hip_event* event);
asc_hip_status asc_sparse_hip_event_elapsed_ms(asc_hip_event* start,
                                               asc_hip_event* stop,
                                               float* out_ms);
asc_hip_status asc_sparse_hip_event_destroy(asc_hip_event* event);

/* ---- staging ------------------------------------------------------------ */

/*
 * Copies a canonical host CSR or CSC matrix into caller-provided device
 * buffers (rows+1 or cols+1 offsets, nnz indices, nnz values) and fills
 * `out` with a trusted device descriptor.  `index_bytes` 8 keeps the
 * reference's int64 indices; 4 is the explicit, validated narrowing clone
 * (fails with ASC_HIP_OVERFLOW if an extent exceeds INT32_MAX).  Replaces
 * CloneCsrErased (providers/cuda.h:212-214; operations.cc:696-800).
 */
asc_hip_status asc_sparse_hip_clone_compressed(
    asc_hip_context* context, const asc_hip_sparse_desc* host_source,
    uint8_t index_bytes, void* device_offsets, void* device_indices,
    void* device_values, asc_hip_sparse_desc* out);

/* Canonical host COO (interleaved int64 coordinates) into device buffers
 * (2*nnz int64 coordinates, nnz values); `out` is a trusted device COO. */
asc_hip_status asc_sparse_hip_clone_coordinate(
    asc_hip_context* context, const asc_hip_sparse_desc* host_source,
    void* device_coordinates, void* device_values, asc_hip_sparse_desc* out);

/* A view of `source` with its value pointer replaced by `values`: same
 * structure pointers, shape, NNZ, memory space and trust.  Host-only; it
 * neither allocates nor copies nor enqueues.  Rejects, as the reference does,
 * a null `values` when nnz > 0 (ASC_HIP_INVALID_ARGUMENT), an address range
 * that wraps (ASC_HIP_OVERFLOW), values overlapping the offsets, indices or
 * coordinates, and values that partly overlap the current values (the same
 * pointer is allowed).  Operations then check, before enqueue, that each
 * operand's whole span lies in one allocation of the context's device, as
 * the runtime reports the allocation's extent (hipMemGetAddressRange, which
 * rounds small requests up to its allocation granularity). 
 * Replaces CompressedSparseView::RebindValues (compressed.h:472-516) and
 * CoordinateView::RebindValues (coordinate.h:533-569). */
asc_hip_status asc_sparse_hip_rebind_values(const asc_hip_sparse_desc* source,
                                            const void* values,
                                            asc_hip_sparse_desc* out);

/* Device-side narrowing of a trusted device matrix's int64 inner indices
 * into `device_indices32` (nnz int32).  `out` shares offsets and values. */
asc_hip_status asc_sparse_hip_narrow_indices(asc_hip_context* context,
                                             const asc_hip_sparse_desc* source,
                                             void* device_indices32,
                                             asc_hip_sparse_desc* out);

/* Device-side value-type change of a trusted matrix (FP64 -> FP32/BF16/F16,
 * round to nearest even) into `device_values`; `out` shares the structure. */
asc_hip_status asc_sparse_hip_convert_values(asc_hip_context* context,
                                             const asc_hip_sparse_desc* source,
                                             uint8_t element_kind,
                                             void* device_values,
                                             asc_hip_sparse_desc* out);

/* The same with every value multiplied by `scale` before rounding; a power
 * of two keeps the scaling exact, so FP16 storage can hold matrices whose
 * magnitudes exceed 65504 (run the SpMV with alpha / scale). */
asc_hip_status asc_sparse_hip_convert_values_scaled(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    uint8_t element_kind, double scale, void* device_values,
    asc_hip_sparse_desc* out);

/* ---- SpMV --------------------------------------------------------------- */

/*
 * Compat CSR SpMV  y = alpha*A*x + beta*y  (y is not read when beta == 0),
 * the device form of asc::Spmv(ctx, alpha, CsrView, x, beta, y)
 * (blas.h:2031-2081; serial algorithm reference_blas.cc:9-41).  Replaces
 * CsrSpmvWorkspaceSizeErased / CsrSpmvErased (providers/cuda.h:220-227).
 * Element kinds: A F64 with x,y F64; A F32/BF16/F16 with x,y F32.
 * `options` may be NULL (ASC_HIP_ALG_AUTO, ASC_HIP_ACC_PLAIN).
 */
asc_hip_status asc_sparse_hip_csr_spmv_workspace_size(
    asc_hip_context* context, const asc_hip_sparse_desc* matrix,
    const asc_hip_vector_desc* x, const asc_hip_vector_desc* y,
    const asc_hip_spmv_options* options, size_t* out_bytes);
asc_hip_status asc_sparse_hip_csr_spmv(
    asc_hip_context* context, double alpha, const asc_hip_sparse_desc* matrix,
    const asc_hip_vector_desc* x, double beta, const asc_hip_vector_desc* y,
    void* workspace, size_t workspace_bytes,
    const asc_hip_spmv_options* options);

/*
 * Allocation-free compat CSR SpMV.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory or pass a workspace (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same operands, options, validation and results as asc_sparse_hip_csr_spmv
 * (it runs the same kernels), but the workspace is the library's own device
 * scratch array of ASC_HIP_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; calls whose algorithm needs more (asc_sparse_hip_csr_spmv_
 * workspace_size) fail with ASC_HIP_UNSUPPORTED before anything is enqueued.
 * Merge path needs about 16 bytes per 512 (rows + nnz) plus 1 MiB, so it
 * serves rows + nnz up to about 10^9; the vector and ordered kernels need
 * none; the stream-staged plan 8 bytes per slice of 16384 entries; the L2 and
 * LDS plans 8 bytes per (panel, listed row).  The call allocates nothing and
 * enqueues no memset.  Every call on one GPU shares that GPU's scratch, so
 * calls that overlap in time on one GPU (two streams, two contexts, two host
 * threads) race: order them on one stream.
 */
#define ASC_HIP_NOALLOC_SCRATCH_BYTES (32u << 20)
asc_hip_status asc_sparse_hip_csr_spmv_noalloc(
    asc_hip_context* context, double alpha, const asc_hip_sparse_desc* matrix,
    const asc_hip_vector_desc* x, double beta, const asc_hip_vector_desc* y,
    const asc_hip_spmv_options* options);

/*
 * Column-panel analysis for a trusted device CSR matrix.  No reference
 * counterpart: it is the analysis step of this library's ASC_HIP_ALG_PANEL
 * SpMV.  With L2 or LDS staging that SpMV reads the caller's CSR arrays in
 * place; with stream staging it reads a copy (below).  Rows with at least
 * `long_row_min` entries are listed; the columns are cut into a multiple of 8
 * panels of about `panel_columns` columns, and the listed rows are summed
 * panel by panel, so that their x gathers hit a resident panel instead of
 * missing to memory: with L2 staging one panel per XCD at a time stays in
 * that XCD's L2; with LDS staging each work-group copies its panel (at most
 * 160 KiB of x) into LDS and walks a chunk of the listed rows.  Shorter rows
 * run in one vector pass.  The
 * analysis buffer (caller-owned device memory, its size from the query)
 * receives the row list and the per-panel split positions; the call waits for
 * its stream.  On success the context's plan cache maps the matrix to
 * ASC_HIP_ALG_PANEL with this buffer: it must stay allocated and unmodified,
 * and the matrix's structure unchanged, until asc_sparse_hip_plan_clear or
 * another plan for the matrix replaces it.  Requires rows and min(nnz, cols)
 * <= UINT32_MAX.  Panel partial sums are added in FP64, so this is a
 * plain-mode (any association) algorithm.
 */
enum {
  ASC_HIP_STAGING_AUTO = 0, /* LDS when a panel fits in 160 KiB of LDS, else L2 */
  ASC_HIP_STAGING_L2 = 1,
  ASC_HIP_STAGING_LDS = 2,  /* ASC_HIP_INVALID_ARGUMENT when a panel does not fit */
  /* Stream: the analysis copies the listed rows' entries panel-major (a
   * 16-bit column within the panel and the value per entry) into the
   * analysis buffer, and the SpMV streams that copy contiguously, gathering x
   * from an LDS copy of each panel (at most 128 KiB of x and 32768 columns
   * per panel).  The copy holds A's values as they were at analysis (or at
   * the last asc_sparse_hip_csr_spmv_refresh_values), so this plan is an
   * explicit opt-in: only ASC_HIP_ALG_PANEL runs it, and ASC_HIP_ALG_AUTO on
   * a matrix with a stream-staged plan runs merge-path on the current values.
   * After the values change in place, refresh the copy before the next
   * ASC_HIP_ALG_PANEL SpMV; a SpMV whose values pointer differs from the
   * copied one (a rebind) fails with ASC_HIP_INVALID_ARGUMENT until a refresh
   * or a new analysis.  Requires nnz <= UINT32_MAX.  The SpMV keeps its
   * (panel, row) partial sums in the analysis buffer, so SpMVs through one
   * analysis must not overlap (one context is never used concurrently).
   * Plain accumulation only (ASC_HIP_ACC_TWOSUM is ASC_HIP_UNSUPPORTED).
   * long_row_min defaults to 1 here: every nonempty row is listed (empty
   * rows, and rows below an explicit long_row_min, run in the vector pass).
   * The size query counts the listed rows and bounds their (panel, row)
   * segments, so it reads the matrix's offsets and waits for the context's
   * stream (a validation-only context returns the bound for every row
   * listed). */
  ASC_HIP_STAGING_STREAM = 3
};

typedef struct asc_hip_panel_options {
  int64_t panel_columns;   /* 0: 2 MiB of x per panel (2^18 FP64 / 2^19 FP32 columns) */
  int64_t long_row_min;    /* 0: 128 */
  uint8_t short_algorithm; /* ASC_HIP_ALG_VECTOR1..VECTOR64, or AUTO */
  uint8_t long_algorithm;  /* ASC_HIP_ALG_VECTOR1..VECTOR64, or AUTO */
  uint8_t staging;         /* ASC_HIP_STAGING_* */
  uint8_t reserved[5];
} asc_hip_panel_options;

asc_hip_status asc_sparse_hip_csr_spmv_analysis_size(
    asc_hip_context* context, const asc_hip_sparse_desc* matrix,
    const asc_hip_panel_options* options, size_t* out_bytes);
asc_hip_status asc_sparse_hip_csr_spmv_analyze(
    asc_hip_context* context, const asc_hip_sparse_desc* matrix,
    const asc_hip_panel_options* options, void* analysis, size_t analysis_bytes);

/*
 * Values-only refresh of a stream-staged panel analysis: enqueues a copy of
 * `matrix`'s current values (same structure and element kind as the analysed
 * matrix; the values pointer may differ, as after a rebind) into the
 * analysis's stream copy, and records that pointer as the copied one.  No
 * reference counterpart: the reference's SpMV always reads the current
 * values (reference_blas.cc:23-33), which ASC_HIP_ALG_AUTO keeps; this call
 * lets an explicit ASC_HIP_ALG_PANEL caller keep the stream plan across
 * value updates (the RebindValues workflow, compressed.h:472-516) at the
 * cost of one pass over the values (nnz * value bytes read and written,
 * 12 bytes per segment of map read), instead of a new analysis.
 * ASC_HIP_INVALID_STATE when the context holds no stream-staged analysis for
 * the matrix.  Does not synchronize.
 */
asc_hip_status asc_sparse_hip_csr_spmv_refresh_values(asc_hip_context* context,
                                                      const asc_hip_sparse_desc* matrix);

/*
 * Standardized SpMV  y += alpha*op(A)*x  for CSR or CSC, real F32/F64
 * (blas.h:1607-1667).  alpha == 0 returns without reading anything.
 * CSR with op N and CSC with op T/H run directly; the other two combinations
 * transpose the structure into the workspace first.  Replaces
 * StandardSpmvErased (providers/cuda.h:240-243).
 */
asc_hip_status asc_sparse_hip_standard_spmv_workspace_size(
    asc_hip_context* context, uint8_t transpose,
    const asc_hip_sparse_desc* matrix, const asc_hip_vector_desc* x,
    const asc_hip_vector_desc* y, const asc_hip_spmv_options* options,
    size_t* out_bytes);
asc_hip_status asc_sparse_hip_standard_spmv(
    asc_hip_context* context, uint8_t transpose, asc_hip_scalar alpha,
    const asc_hip_sparse_desc* matrix, const asc_hip_vector_desc* x,
    const asc_hip_vector_desc* y, void* workspace, size_t workspace_bytes,
    const asc_hip_spmv_options* options);

/* ---- sparse Level-1 ------------------------------------------------------ */

/* Indexed (sparse) vector, after IndexedVectorDescriptor
 * (providers/cuda.h:175-184) and SparseBlasIndexedVectorView
 * (blas.h:526-789): `nnz` int64 indices, strictly increasing in
 * [0, dense_extent), and `nnz` values. */
typedef struct asc_hip_indexed_desc {
  const int64_t* indices;
  const void* values;
  int64_t nnz;
  int64_t dense_extent;
  uint8_t element_kind; /* ASC_HIP_F32 / F64 */
  uint8_t memory_space;
  uint8_t writable;
  uint8_t trusted; /* canonical structure vouched for by this library */
  uint8_t reserved[4];
} asc_hip_indexed_desc;

/* internal_sparse_cuda::StandardOperation (providers/cuda.h:111-117). */
enum {
  ASC_HIP_LEVEL1_DOT = 0,
  ASC_HIP_LEVEL1_AXPY = 1,
  ASC_HIP_LEVEL1_GATHER = 2,
  ASC_HIP_LEVEL1_GATHER_ZERO = 3,
  ASC_HIP_LEVEL1_SCATTER = 4
};
/* asc::SparseBlasConjugation (blas.h:74-77); the identity for real data. */
enum { ASC_HIP_UNCONJUGATED = 0, ASC_HIP_CONJUGATED = 1 };

/* Copies a host indexed vector, checked canonical (blas.h:606-620:
 * ASC_HIP_INDEX otherwise), into device buffers (nnz int64 indices, nnz
 * values); `out` is trusted.  Replaces CloneIndexedErased
 * (providers/cuda.h:216-218). */
asc_hip_status asc_sparse_hip_clone_indexed(asc_hip_context* context,
                                            const asc_hip_indexed_desc* host_source,
                                            void* device_indices, void* device_values,
                                            asc_hip_indexed_desc* out);

/*
 * SparseDot / SparseAxpy / SparseGather / SparseGatherZero / SparseScatter
 * (blas.h:1330-1582) on a trusted device indexed vector and a device dense
 * vector of extent dense_extent.  Replaces StandardLevel1Erased
 * (providers/cuda.h:234-238).
 *  - DOT writes sum(v[p] * x[idx[p]]) to `result` (one element, device).
 *    `options->algorithm` ASC_HIP_ALG_ORDERED sums left to right from +0
 *    without FMA (bit-exact with the serial reference); AUTO (or NULL
 *    options) uses a deterministic parallel tree.
 *  - AXPY: x[idx[p]] += alpha * v[p]; alpha == 0 returns without reading.
 *  - GATHER: v[p] = x[idx[p]]; GATHER_ZERO also sets x[idx[p]] = 0.
 *  - SCATTER: x[idx[p]] = v[p].
 * Writable operands must not overlap the others.
 */
asc_hip_status asc_sparse_hip_standard_level1(asc_hip_context* context, uint8_t operation,
                                              uint8_t conjugation, asc_hip_scalar alpha,
                                              const asc_hip_indexed_desc* sparse,
                                              const asc_hip_vector_desc* dense,
                                              const asc_hip_vector_desc* result,
                                              const asc_hip_spmv_options* options);

/* ---- triangular solve ---------------------------------------------------- */

/* asc::SparseBlasTriangle / SparseBlasDiagonal (blas.h:117-134). */
enum { ASC_HIP_UPPER = 0, ASC_HIP_LOWER = 1 };
enum { ASC_HIP_NON_UNIT = 0, ASC_HIP_UNIT = 1 };

/* What an analysis describes; the solve checks it against its arguments. */
typedef struct asc_hip_triangular_info {
  int64_t rows;
  int64_t nnz;
  int64_t levels;          /* solve launches: rows of one level run in parallel */
  int64_t max_level_rows;
  const void* offsets;     /* the analysed matrix's arrays */
  const void* indices;
  const void* values;
  uint8_t transpose;
  uint8_t triangle;
  uint8_t diagonal;
  uint8_t format;
  uint8_t reserved[4];
} asc_hip_triangular_info;

/*
 * Level-scheduled sparse triangular solve, the device form of
 * SparseTriangularSolve / SparseTriangularSolveMultiple (blas.h:1760-1880);
 * replaces StandardTriangularSolveErased (providers/cuda.h:250-253).  No
 * reference counterpart for the split into analysis and solve.
 *  - analyze (synchronous) checks what SparseBlasTriangularView::Create
 *    checks (blas.h:1090-1160: square, every entry inside the declared
 *    triangle, a stored nonzero diagonal when non-unit; ASC_HIP_SHAPE /
 *    ASC_HIP_INVALID_ARGUMENT otherwise), then levels the rows of op(A) into
 *    the caller's analysis buffer (with a transposed copy of A for CSR with
 *    op T/H and CSC with op N).  The buffer must stay unmodified while it is
 *    used.
 *  - solve overwrites the n x k right-hand sides with
 *    alpha * op(A)^-1 * B: one launch per level, each (row, column) a serial
 *    chain over the row's stored entries in column order without FMA.  That
 *    equals the reference's dense loop bit for bit while the solution stays
 *    finite.  alpha == 0 zeroes the right-hand sides.
 */
asc_hip_status asc_sparse_hip_triangular_analysis_size(asc_hip_context* context,
                                                       uint8_t transpose,
                                                       const asc_hip_sparse_desc* matrix,
                                                       size_t* out_bytes);
asc_hip_status asc_sparse_hip_triangular_analyze(asc_hip_context* context, uint8_t transpose,
                                                 uint8_t triangle, uint8_t diagonal,
                                                 const asc_hip_sparse_desc* matrix,
                                                 void* analysis, size_t analysis_bytes,
                                                 asc_hip_triangular_info* out_info);
asc_hip_status asc_sparse_hip_triangular_solve(asc_hip_context* context, uint8_t transpose,
                                               asc_hip_scalar alpha,
                                               const asc_hip_sparse_desc* matrix,
                                               const asc_hip_matrix_desc* right_hand_sides,
                                               const void* analysis, size_t analysis_bytes,
                                               const asc_hip_triangular_info* info);

/* ---- SpMM --------------------------------------------------------------- */

/*
 * Standardized SpMM  C += alpha*op(A)*B  against a dense block in either
 * layout (blas.h:1692-1750).  Replaces StandardSpmmErased
 * (providers/cuda.h:245-248); `options` (may be NULL) picks the form:
 *  - ASC_HIP_ALG_ORDERED: every C(i,j) is accumulated in stored order
 *    without FMA, so the result equals the serial reference bit for bit.
 *    On the plain form's operands (below) it gathers B in the same column
 *    passes when the workspace holds their split table (one pass
 *    otherwise), and when the context holds a stream-staged analysis of
 *    this CSR matrix listing only long rows (as for AUTO, below) it takes
 *    those rows panel by panel; neither changes a bit of the result;
 *  - ASC_HIP_ALG_AUTO (NULL): the plain form (below ASC_HIP_ALG_PANEL for
 *    when AUTO takes the panel form).  For row-major B and C with
 *    k <= 32, k and both leading dimensions multiples of 4 and B and C
 *    aligned to 4 elements, each C(i,j) sums its terms in registers with
 *    fused multiply-adds and is updated once, within
 *    4*len_i*eps*|alpha|*(|A||B|)(i,j) of the serial result; B is gathered
 *    in slices of 256 MiB, one pass over the matrix per slice, when the
 *    workspace holds the passes' split table ((passes-1)*rows uint32; a
 *    smaller workspace runs one pass).  Other shapes run the ordered form.
 *  - ASC_HIP_ALG_PANEL: the plain form with a stream-staged column-panel
 *    analysis of this CSR matrix (asc_sparse_hip_csr_spmv_analyze with
 *    ASC_HIP_STAGING_STREAM; ASC_HIP_INVALID_ARGUMENT without one), op(A) = A
 *    and the plain form's operands (else ASC_HIP_UNSUPPORTED).  The rows the
 *    analysis lists go segment by segment (a row's entries in one panel),
 *    gathering B from the panel's slice, which one XCD's L2 keeps; their
 *    partial sums, kept in the workspace (32 columns * element bytes per
 *    segment, plus 4 per listed row), are added in panel order, so the
 *    result is within the same bound as the plain form and repeats bit for
 *    bit.  The other rows run the plain kernel and equal its result.  The
 *    SpMM reads the matrix's current indices and values, not the
 *    analysis's copy.  ASC_HIP_ALG_AUTO takes this form by itself when the
 *    context holds such an analysis whose listed rows have at least 8
 *    entries per panel (long_row_min >= 8 * panels) and the workspace holds
 *    the partial sums (a smaller workspace runs the plain kernel over every
 *    row); an analysis listing every row, the stream form's default, leaves
 *    AUTO on the plain kernel.
 *    Results repeat bit for bit from run to run in every form.
 * `options->accumulate` must be ASC_HIP_ACC_PLAIN; other algorithms are
 * ASC_HIP_INVALID_ARGUMENT.
 */
asc_hip_status asc_sparse_hip_standard_spmm_workspace_size(
    asc_hip_context* context, uint8_t transpose,
    const asc_hip_sparse_desc* matrix, const asc_hip_matrix_desc* b,
    const asc_hip_matrix_desc* c, const asc_hip_spmv_options* options,
    size_t* out_bytes);
asc_hip_status asc_sparse_hip_standard_spmm(
    asc_hip_context* context, uint8_t transpose, asc_hip_scalar alpha,
    const asc_hip_sparse_desc* matrix, const asc_hip_matrix_desc* b,
    const asc_hip_matrix_desc* c, void* workspace, size_t workspace_bytes,
    const asc_hip_spmv_options* options);

/* ---- conversions --------------------------------------------------------- */

/*
 * Canonical device COO -> CSR (out_format ASC_HIP_CSR) or CSC
 * (ASC_HIP_CSC), bit-exact with ConvertToCsr / ConvertToCsc
 * (compressed.h:1675-1708 -> FromCoordinate :1082-1131).  Outputs are
 * caller-provided device buffers (outer+1 offsets int64, nnz indices of
 * `index_bytes`, nnz values); `out` receives a trusted descriptor.  The
 * reference provider has no device conversion (docs/modules/sparse.md:592).
 */
asc_hip_status asc_sparse_hip_convert_coordinate_workspace_size(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    uint8_t out_format, uint8_t index_bytes, size_t* out_bytes);
asc_hip_status asc_sparse_hip_convert_coordinate(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    uint8_t out_format, uint8_t index_bytes, void* out_offsets,
    void* out_indices, void* out_values, asc_hip_sparse_desc* out,
    void* workspace, size_t workspace_bytes);

/*
 * Trusted device CSR -> CSC or CSC -> CSR, bit-exact with
 * ConvertToCsc(ctx, CsrView) / ConvertToCsr(ctx, CscView)
 * (compressed.h:1727-1760 -> FromCompressed :1151-1221), by a stable device
 * radix sort on the inner index.
 */
asc_hip_status asc_sparse_hip_convert_compressed_workspace_size(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    size_t* out_bytes);
asc_hip_status asc_sparse_hip_convert_compressed(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    void* out_offsets, void* out_indices, void* out_values,
    asc_hip_sparse_desc* out, void* workspace, size_t workspace_bytes);

/*
 * Trusted device CSR or CSC -> canonical COO (row-major, interleaved int64
 * coordinates; values moved bit for bit), the device form of
 * ConvertToCoordinate (compressed.h:1781-1837).  CSC sources are transposed
 * in the workspace first (zero workspace for CSR).
 */
asc_hip_status asc_sparse_hip_convert_to_coordinate_workspace_size(
    asc_hip_context* context, const asc_hip_sparse_desc* source, size_t* out_bytes);
asc_hip_status asc_sparse_hip_convert_to_coordinate(
    asc_hip_context* context, const asc_hip_sparse_desc* source,
    void* device_coordinates, void* device_values, asc_hip_sparse_desc* out,
    void* workspace, size_t workspace_bytes);

/*
 * Device CoordinateBuilder::Finalize (coordinate.h:1300-1497): raw device
 * triplets (a COO descriptor of interleaved int64 (row, col) coordinates and
 * F32/F64 values, in insertion order; `trusted` not required, nnz may exceed
 * rows*cols through duplicates) -> canonical COO in caller-provided buffers
 * sized for `nnz` entries (2*nnz int64 coordinates, nnz values).
 *  - Coordinates outside the shape fail with ASC_HIP_INDEX.
 *  - ASC_HIP_DUPLICATE_REJECT fails with ASC_HIP_INVALID_ARGUMENT if any
 *    coordinate repeats; ASC_HIP_DUPLICATE_SUM sums each run of equal
 *    coordinates left to right in insertion order (no FMA).
 *  - ASC_HIP_ZERO_DROP then drops entries equal to zero (-0.0 included,
 *    NaN kept); ASC_HIP_ZERO_KEEP keeps them.
 * On failure the outputs are untouched.  The call waits for its stream and
 * returns the finalized descriptor (its nnz is the entry count) in `out`.
 * rows, cols and nnz must not exceed UINT32_MAX.  The reference builder is
 * host-only; this is its device form.
 */
asc_hip_status asc_sparse_hip_coo_finalize_workspace_size(
    asc_hip_context* context, const asc_hip_sparse_desc* triplets, size_t* out_bytes);
asc_hip_status asc_sparse_hip_coo_finalize(
    asc_hip_context* context, const asc_hip_sparse_desc* triplets,
    uint8_t duplicate_policy, uint8_t zero_policy, void* out_coordinates,
    void* out_values, asc_hip_sparse_desc* out, void* workspace,
    size_t workspace_bytes);

/* ---- batched SpMV --------------------------------------------------------- */

/*
 * Thousands of independent compat SpMVs  y_i = alpha_i*A_i*x_i + beta_i*y_i
 * in one persistent launch.  No reference counterpart: every entry follows
 * the contract of asc::Spmv(ctx, alpha, CsrView, x, beta, y)
 * (blas.h:2031-2081), entries share element kind (F32/F64) and index width,
 * and no output may overlap any input or other output.
 *
 * A plan is a work list of (matrix, row range, lanes per row) items in
 * caller-owned device memory.  Lanes per row come from `classes`: class c
 * holds matrices whose average row length lies in [2^c, 2^(c+1)) (class 7
 * is open-ended); NULL selects {1,2,4,8,16,32,64,64}.  Lanes 0 selects
 * CSR-stream items: a work-group stages an item's products (up to 2048
 * entries) in LDS with coalesced loads and sums each row from there; an
 * item with more entries runs one lane per row.  The plan stores raw
 * pointers: entries' buffers must outlive it.
 */
typedef struct asc_hip_batch_entry {
  asc_hip_sparse_desc matrix;
  asc_hip_vector_desc x;
  asc_hip_vector_desc y;
  double alpha;
  double beta;
} asc_hip_batch_entry;

typedef struct asc_hip_batch_classes {
  uint8_t lanes[8];
} asc_hip_batch_classes;

typedef struct asc_hip_batch_plan {
  void* matrix_table; /* device: per-matrix descriptors */
  void* item_table;   /* device: work items */
  int64_t matrices;
  int64_t items;
  int64_t grid; /* persistent workgroups */
  int32_t device;
  uint8_t element_kind;
  uint8_t index_bytes;
  uint8_t reserved[2];
} asc_hip_batch_plan;

typedef struct asc_hip_graph asc_hip_graph;

asc_hip_status asc_sparse_hip_batch_plan_size(asc_hip_context* context,
                                              int64_t count,
                                              const asc_hip_batch_entry* entries,
                                              size_t* out_bytes);
asc_hip_status asc_sparse_hip_batch_plan_build(
    asc_hip_context* context, int64_t count, const asc_hip_batch_entry* entries,
    const asc_hip_batch_classes* classes, void* plan_memory, size_t plan_bytes,
    asc_hip_batch_plan* out);
asc_hip_status asc_sparse_hip_batch_spmv(asc_hip_context* context,
                                         const asc_hip_batch_plan* plan);

/* Captures the plan's launch into a hipGraph (one kernel node, no memset
 * nodes) for replay on the context stream. */
asc_hip_status asc_sparse_hip_batch_graph_create(asc_hip_context* context,
                                                 const asc_hip_batch_plan* plan,
                                                 asc_hip_graph** out);
asc_hip_status asc_sparse_hip_graph_launch(asc_hip_context* context,
                                           asc_hip_graph* graph);
asc_hip_status asc_sparse_hip_graph_destroy(asc_hip_graph* graph);

/* Exhaustive per-class autotuner: times every lanes-per-row choice (1..64,
 * then 0 for CSR-stream) on each class's matrices with `reps` launches and
 * writes the fastest to `out`; `class_ms` (8 x 8 floats, may be NULL)
 * receives the timings in that order (-1 for empty classes).  Uses
 * `plan_memory` as scratch; outputs are written. */
asc_hip_status asc_sparse_hip_batch_tune(asc_hip_context* context, int64_t count,
                                         const asc_hip_batch_entry* entries,
                                         void* plan_memory, size_t plan_bytes,
                                         int32_t reps, asc_hip_batch_classes* out,
                                         float* class_ms);

/* ---- variant selection: autotuner, plan cache, features, learned MLP ------ */

/*
 * No reference counterpart: the reference provider always runs one library
 * algorithm (src/sparse/cuda/operations.cc:1004-1040).  The context keeps a
 * small plan cache (matrix identity -> algorithm) consulted by
 * ASC_HIP_ALG_AUTO; autotuning and selection fill it.
 */
asc_hip_status asc_sparse_hip_plan_set(asc_hip_context* context,
                                       const asc_hip_sparse_desc* matrix,
                                       uint8_t algorithm);
asc_hip_status asc_sparse_hip_plan_get(asc_hip_context* context,
                                       const asc_hip_sparse_desc* matrix,
                                       uint8_t* out_algorithm);
asc_hip_status asc_sparse_hip_plan_clear(asc_hip_context* context);

/* Exhaustive autotuner: times every tunable variant (ASC_HIP_ALG_MERGE ..
 * ASC_HIP_ALG_VECTOR64) as y = A*x with `reps` launches after a warm-up,
 * writes per-algorithm milliseconds into times_ms[ASC_HIP_ALG_VECTOR64 + 1]
 * (indexed by algorithm id; may be NULL), stores the fastest in the plan
 * cache and returns it.  y is overwritten; the workspace must hold what
 * the merge variant needs. */
asc_hip_status asc_sparse_hip_csr_spmv_autotune(
    asc_hip_context* context, const asc_hip_sparse_desc* matrix,
    const asc_hip_vector_desc* x, const asc_hip_vector_desc* y, void* workspace,
    size_t workspace_bytes, int32_t reps, float* times_ms, uint8_t* out_algorithm);

/* Sixteen sparsity features of a trusted device CSR matrix (FP32, device
 * output), computed in one deterministic pass. */
asc_hip_status asc_sparse_hip_csr_features_workspace_size(asc_hip_context* context,
                                                          size_t* out_bytes);
asc_hip_status asc_sparse_hip_csr_features(asc_hip_context* context,
                                           const asc_hip_sparse_desc* matrix,
                                           float* device_features16, void* workspace,
                                           size_t workspace_bytes);

/* The selector MLP 16 -> 64 -> 64 -> k (ReLU), FP32, on device buffers:
 * params (see asc_sparse_hip_mlp_sizes), x [batch][16], y / target
 * [batch][k].  The gradient is of the mean squared error over batch*k
 * outputs; `loss` (device float) receives that error. */
asc_hip_status asc_sparse_hip_mlp_sizes(int32_t k, int64_t batch,
                                        size_t* param_floats, size_t* scratch_bytes);
asc_hip_status asc_sparse_hip_mlp_forward(asc_hip_context* context, int32_t k,
                                          const float* params, int64_t batch,
                                          const float* x, float* y, void* scratch,
                                          size_t scratch_bytes);
asc_hip_status asc_sparse_hip_mlp_gradient(asc_hip_context* context, int32_t k,
                                           const float* params, int64_t batch,
                                           const float* x, const float* target,
                                           float* grads, float* loss, void* scratch,
                                           size_t scratch_bytes);
asc_hip_status asc_sparse_hip_mlp_adam(asc_hip_context* context, float* params,
                                       const float* grads, float* m, float* v,
                                       int64_t n, float lr, float beta1, float beta2,
                                       float eps, int64_t step);
/* out[i][f] = (x[i][f] - mean[f]) / stddev[f] for n rows of 16 features. */
asc_hip_status asc_sparse_hip_normalize_features(asc_hip_context* context,
                                                 const float* x, const float* mean,
                                                 const float* stddev, int64_t n,
                                                 float* out);

/* A trained selector: device_model = [mean 16][stddev 16][MLP params].
 * `outputs` is 8 (the direct kernels) or 11 (the direct kernels and the
 * column-panel plans).  Output j predicts the log of the variant's cost
 * over the best variant's:
 *   j = 0..7:  SpMV time of ASC_HIP_ALG_MERGE + j (merge, vec1 .. vec64);
 *   j = 8..10: ASC_HIP_ALG_PANEL with staging ASC_HIP_STAGING_L2, _LDS,
 *              _STREAM (default panel widths): SpMV time plus the analysis
 *              time spread over ASC_HIP_SELECT_AMORTIZE SpMVs.
 * Selection takes the smallest prediction unless unsure -- some normalized
 * feature outside [-clip, clip], or the merge-path prediction within
 * `margin` of the best -- in which case it falls back to merge-path. */
#define ASC_HIP_SELECT_AMORTIZE 100
#define ASC_HIP_SELECT_DIRECT_OUTPUTS 8
#define ASC_HIP_SELECT_PLAN_OUTPUTS 11
typedef struct asc_hip_selector {
  const float* device_model;
  int32_t outputs;
  float margin;
  float clip;
  int32_t reserved;
} asc_hip_selector;

/* The choice.  A direct pick (algorithm ASC_HIP_ALG_MERGE..VECTOR64) is
 * stored in the plan cache, so ASC_HIP_ALG_AUTO runs it.  A plan pick
 * (algorithm ASC_HIP_ALG_PANEL, `staging` set) needs caller memory: the
 * caller sizes and runs asc_sparse_hip_csr_spmv_analyze with that staging,
 * which registers the plan; until then the plan cache holds
 * `direct_algorithm`, the best-predicted direct kernel (merge-path when
 * unsure). */
typedef struct asc_hip_selection {
  uint8_t algorithm;
  uint8_t staging;          /* ASC_HIP_STAGING_* for ASC_HIP_ALG_PANEL, else 0 */
  uint8_t direct_algorithm; /* what the plan cache holds after the call */
  uint8_t fell_back;        /* 1 when unsure: merge-path */
  uint8_t reserved[4];
} asc_hip_selection;

asc_hip_status asc_sparse_hip_select_workspace_size(asc_hip_context* context,
                                                    const asc_hip_selector* selector,
                                                    size_t* out_bytes);
asc_hip_status asc_sparse_hip_select(asc_hip_context* context,
                                     const asc_hip_sparse_desc* matrix,
                                     const asc_hip_selector* selector, void* workspace,
                                     size_t workspace_bytes, asc_hip_selection* out,
                                     float* out_scores);

/* ---- host-forced fault injection (tests) -------------------------------- */

/* Makes the `countdown`-th next allocation (0 = the very next) or event
 * record fail as if HIP had failed; -1 disarms.  Mirrors the reference's
 * TrackingMemoryResource::FailOnAllocation (tests/sparse/test_resources.h:
 * 47-53) and its CUDA runtime test hooks (src/core/cuda/runtime.cc:53-70). */
/* ASC_HIP_INJECT_STREAM_STALL (no reference counterpart) makes the matching
 * stream-staged panel SpMV run with its inter-wave waits bounded at zero and
 * no writer wave, so that a wait gives up as a protocol fault would: the
 * kernel ends normally, its outputs are not valid, and the next
 * asc_sparse_hip_context_synchronize fails with ASC_HIP_INTERNAL. */
enum { ASC_HIP_INJECT_ALLOCATION = 0, ASC_HIP_INJECT_EVENT = 1, ASC_HIP_INJECT_STREAM_STALL = 2 };
asc_hip_status asc_sparse_hip_test_inject(uint32_t what, int32_t countdown);

#ifdef __cplusplus
}
#endif

#endif /* ASC_SPARSE_HIP_H_ */
