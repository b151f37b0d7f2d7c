// Host-side launch interface between the C ABI (csrc/abi) and the kernels.
// Every launcher assumes its arguments were validated by the ABI layer
// (shapes, element kinds, trust, workspace size) and only enqueues work.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace asc_hip {

enum class Elem : uint8_t { F32 = 0, F64 = 1, BF16 = 4, F16 = 5 };

// A compressed operand seen as "rows" of the kernel: outer = row-like
// dimension.  Offsets are int64; indices int32 or int64.
struct Compressed {
  const int64_t* offsets;
  const void* indices;
  const void* values;
  int64_t outer;
  int64_t inner;
  int64_t nnz;
  int index_bytes;
  Elem elem;
};

enum class SpmvAlg : uint8_t {
  Ordered = 1, Merge = 2,
  Vec1 = 3, Vec2 = 4, Vec4 = 5, Vec8 = 6, Vec16 = 7, Vec32 = 8, Vec64 = 9,
  Panel = 10  // column-panel plan for long rows (see PanelPlan)
};

enum class SpmvForm : uint8_t {
  Compat = 0,    // y = alpha*sum + beta*y (y unread when beta == 0)
  Standard = 1,  // y += alpha*v*x per entry (ordered) / y += alpha*sum (plain)
};

// Column-panel plan (built by panel_analysis_launch): rows with at least
// `long_min` entries are listed in `long_rows` (ascending); the columns are cut
// into `panels` ranges of `panel_cols` (a multiple of 8 ranges, one per XCD at
// a time).  For listed row i, row_begin[i] is its first entry and
// splits[q*nlong + i] (q = 0..panels-1) the end of its panel-q entries
// relative to row_begin[i] (the last panel's end is the row length).  Shorter
// rows run through vector<short_alg> in one pass.
struct PanelPlan {
  const uint32_t* long_rows = nullptr;
  const int64_t* row_begin = nullptr;
  const uint32_t* splits = nullptr;
  int64_t nlong = 0;
  int64_t panels = 8;
  int64_t panel_cols = 0;
  int64_t long_min = INT64_MAX;
  SpmvAlg short_alg = SpmvAlg::Vec4;  // Vec1..Vec64
  SpmvAlg long_alg = SpmvAlg::Vec32;  // Vec1..Vec64
  bool lds = false;     // stage each x panel in LDS (panel_cols * x bytes <= kLdsPanelBytes)
  int64_t chunks = 1;   // LDS form: row chunks per panel (work-groups = panels * chunks)
  // Stream form (built by panel_stream_analysis_launch): the listed rows'
  // entries copied panel-major, so the SpMV streams them contiguously and
  // gathers x from an LDS copy of the panel.  A listed row's entries in one
  // panel form a segment; segments are numbered panel-major (rows ascending
  // inside a panel).  Panel q's entries are padded to a multiple of
  // kStreamTile; scol holds the column within the panel, with kStreamHead set
  // on each segment's first entry, sval the value.  A slice (kStreamSlice
  // entries at most, one wave) is described by slice_info[4k..4k+3] =
  // {start, end, segment open before the first entry, panel | kStreamClosed
  // when its last segment ends inside it}; wg_table[2w..] = {first slice,
  // slice count} of work-group w (one panel, XCD-ordered).  spartial[ordinal]
  // holds each segment's sum; seg_lrow[ordinal] its listed row modulo
  // kStreamGroupRows, and ctable[q*(groups+1) + g] the first ordinal of panel
  // q whose listed row is at least g*kStreamGroupRows.
  bool stream = false;
  const uint16_t* scol = nullptr;
  const void* sval = nullptr;
  const uint16_t* seg_lrow = nullptr;
  const uint32_t* ctable = nullptr;
  int64_t groups = 0;
  const uint32_t* slice_info = nullptr;
  const uint32_t* wg_table = nullptr;
  double* spartial = nullptr;
  int64_t slices = 0;
  int64_t wgs = 0;
  const void* values_src = nullptr;  // the values the copy was taken from
  size_t stream_bytes = 0;           // bytes of the stream arrays (overlap checks)
  const void* stream_base = nullptr;
  // Per segment in ordinal order: its first entry in the CSR arrays, its
  // first (padded) stream position and its length, so a values-only refresh
  // can copy new values into the stream without redoing the analysis.
  const uint32_t* seg_src = nullptr;
  const uint32_t* seg_dst = nullptr;
  const uint32_t* seg_len = nullptr;
  int64_t nseg = 0;
};

// LDS bytes one work-group stages an x panel in (gfx950: 160 KiB per CU).
constexpr int64_t kLdsPanelBytes = int64_t(160) << 10;
// Stream form geometry: entries per lane per step, entries per wave step,
// entries per slice (one wave's share), slices (streaming waves) per
// work-group, and the work-group: those waves plus one writer wave.
constexpr int kStreamEpl = 8;
constexpr int kStreamTile = 64 * kStreamEpl;
constexpr int kStreamSlice = 32 * kStreamTile;
constexpr int kStreamWaves = 7;
constexpr int kStreamBlock = 64 * (kStreamWaves + 1);
// Step buffers a streaming wave rotates through: kStreamDepth - 1 steps stay
// in flight while one is reduced.
constexpr int kStreamDepth = 4;
// Segment sums a wave keeps in LDS for the writer wave (a power of two), one
// spare slot per lane beside them, and the LDS left for the x panel.
constexpr int kStreamRing = 512;
// Bound on each wait between the stream kernel's waves (sleeps of about
// 64 clocks; about half a second): a wave that reaches it stops, records
// kFaultStreamWait in the context's fault word and leaves, so a protocol
// error ends the kernel instead of hanging it.  0 is a test setting: the
// writer wave leaves at once and every wait gives up at its first check.
constexpr uint32_t kStreamSpinLimit = 1u << 24;
constexpr uint32_t kFaultStreamWait = 1u;
constexpr int64_t kStreamPanelBytes = int64_t(128) << 10;
static_assert(kStreamPanelBytes + int64_t(kStreamWaves) * ((kStreamRing + 64) * 8 + 3 * 4) <=
                  kLdsPanelBytes,
              "x panel, segment rings and ring counters share one work-group's LDS");
constexpr uint16_t kStreamHead = 0x8000;
constexpr uint32_t kStreamClosed = 0x80000000u;
// Listed rows per combine wave (one FP64 LDS accumulator each), and combine
// waves per 1024-thread work-group.
constexpr int kStreamGroupRows = 1024;
constexpr int kCombineWaves = 16;
// Combine work-groups wanted at least (one per CU of an MI355X).
constexpr int64_t kCombineTargetBlocks = 256;
// Widest stream panel: the column field keeps 15 bits beside the head flag.
constexpr int64_t kStreamMaxPanelCols = 32768;

struct SpmvArgs {
  Compressed a;
  const void* x;
  int64_t incx;
  void* y;
  int64_t incy;
  double alpha;
  double beta;
  SpmvForm form;
  SpmvAlg alg;
  bool twosum;  // reduced-precision compensated accumulation
  PanelPlan panel;  // used when alg == Panel
  // Stream form: where a wave that gives up waiting records it (may be
  // null), and the wait bound.
  uint32_t* fault = nullptr;
  uint32_t spin_limit = kStreamSpinLimit;
};

// A stream beside the main one and two events: `fork` recorded on the main
// stream before side work starts, `join` recorded on the side stream after it.
struct SideStream {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  hipEvent_t join = nullptr;
};

// Workspace bytes the algorithm needs for this operand (0 for most).
size_t spmv_workspace_bytes(const SpmvArgs& args);
// With `side`, the stream-form panel plan runs its rows below long_min (if
// any) on the side stream, beside the stream kernels; the main stream waits
// for them before it returns.
hipError_t spmv_launch(const SpmvArgs& args, void* workspace,
                       hipStream_t stream, const SideStream* side = nullptr);

// Builds a PanelPlan's row list, row starts and split table for a canonical
// CSR operand (sorted columns per row): `long_rows` holds `outer` uint32,
// `row_begin` `outer` int64 and `splits` panels*outer uint32 entries;
// `scratch` holds panel_analysis_scratch_bytes(outer).  Waits for the stream
// to read back the number of listed rows.
size_t panel_analysis_scratch_bytes(int64_t outer);
hipError_t panel_analysis_launch(const Compressed& a, int64_t panel_cols, int64_t panels,
                                 int64_t long_min, uint32_t* long_rows, int64_t* row_begin,
                                 uint32_t* splits, void* scratch, int64_t* out_nlong,
                                 hipStream_t stream);

// Stream-form layout for a matrix whose listed rows (nlong of them, nnz_long
// entries, at most segs_cap (panel, row) segments) are known: the byte
// offsets of every array in the analysis buffer.
struct StreamLayout {
  int64_t nlong = 0, nnz_long = 0, panels = 0, groups = 0;
  int64_t entries_cap = 0, slices_cap = 0, segs_cap = 0;
  size_t long_rows = 0, partial = 0, scol = 0, sval = 0, seg_lrow = 0, ctable = 0,
         slice_info = 0, wg_table = 0, seg_src = 0, seg_dst = 0, seg_len = 0;  // kept
  size_t row_begin = 0, cnt = 0, sbase = 0, sstart = 0, spanel = 0, srow = 0, slen = 0, sorder = 0,
         lens_s = 0, lrow_s = 0, dest = 0, pseg = 0, small = 0, slice_stage = 0, scan = 0,
         sort = 0, list_scratch = 0;  // analysis only
  size_t kept_bytes = 0, total = 0;
};
StreamLayout stream_layout(int64_t rows, int64_t nlong, int64_t nnz_long, int64_t segs_cap,
                           int64_t panels, size_t value_bytes);
// Builds the stream form into `base` (laid out by `L`); fills the stream
// fields of `pp` (whose panels/panel_cols/long_min are set).  Synchronous.
hipError_t panel_stream_analysis_launch(const Compressed& a, const StreamLayout& L, void* base,
                                        PanelPlan* pp, hipStream_t stream);
// Copies `values` (the matrix's current values, same element kind as the
// analysed ones) into the stream of a stream-form plan, segment by segment.
hipError_t panel_stream_refresh_launch(const PanelPlan& pp, const void* values,
                                       size_t value_bytes, hipStream_t stream);

// Stable sort of the indices 0..n-1 by uint32 key (< key_bound): out_index
// gets the sorted order, out_key_offsets[k] (k = 0..key_bound) the first
// sorted position with key >= k.  `workspace` holds
// sort_by_key_workspace_bytes(n, key_bound).
size_t sort_by_key_workspace_bytes(int64_t n, int64_t key_bound);
hipError_t sort_by_key_launch(const uint32_t* keys, int64_t n, int64_t key_bound,
                              uint32_t* out_index, int64_t* out_key_offsets, void* workspace,
                              hipStream_t stream);

// ---- triangular solve (level-scheduled) ------------------------------------
// The row view `rows` is op(A) in CSR form (outer = n rows of op(A)).
// Analysis: validates the stored matrix `stored` against its declared
// triangle (stored_inner_le_outer: every inner index <= its outer index) and,
// when !unit, a stored nonzero diagonal (returns *failure 1 / 2), then levels
// the rows of the row view (`lower`: the solve runs forward) and sorts them
// by level.  Synchronous.
struct TrsvLevels {
  int64_t levels = 0;
  int64_t max_level_rows = 0;
  // level_ptr on the host (levels + 1 entries) when known: the solve then
  // runs each stretch of small levels in one work-group launch.
  std::vector<int64_t> host_level_ptr;
};
size_t trsv_analysis_scratch_bytes(int64_t n);
hipError_t trsv_analysis_launch(const Compressed& stored, bool stored_inner_le_outer, bool unit,
                                const Compressed& rows, bool lower, int64_t* level_ptr,
                                uint32_t* row_order, void* scratch, TrsvLevels* out,
                                int* failure, hipStream_t stream);
// Solve in place on X (n x k, element (r, j) at x[r*rs + j*cs]); alpha == 0
// is handled by the caller.
hipError_t trsv_solve_launch(const Compressed& rows, bool lower, bool unit, double alpha,
                             const int64_t* level_ptr, const uint32_t* row_order,
                             const TrsvLevels& levels, void* x, int64_t rs, int64_t cs,
                             int64_t k, hipStream_t stream);
hipError_t dense_zero_launch(Elem elem, void* x, int64_t rows, int64_t cols, int64_t rs,
                             int64_t cs, hipStream_t stream);

// Exclusive prefix sum of n uint32 values; `scratch` holds
// scan_u32_scratch_bytes(n).
size_t scan_u32_scratch_bytes(int64_t n);
hipError_t scan_u32_exclusive(const uint32_t* in, uint32_t* out, int64_t n, void* scratch,
                              hipStream_t stream);

struct SpmmArgs {
  Compressed a;
  const void* b;
  int64_t ldb;
  bool b_row_major;
  void* c;
  int64_t ldc;
  bool c_row_major;
  int64_t k;  // right-hand sides
  double alpha;
  // true: every C(i,j) receives its terms in stored order without FMA (bit-
  // exact with the serial reference); false: the plain kernel where it
  // applies (row-major B and C, k <= 32, k/ldb/ldc multiples of 4, 4-element
  // aligned B and C), fused multiply-adds in a fixed per-row order.
  bool ordered = true;
  // Columns of op(A) (rows of B) per pass; 0 or >= inner: one pass.  Each
  // pass gathers only from its slice of B, which stays memory-side cached.
  int64_t panel_cols = 0;
  // Plain form with a stream-staged column-panel plan of op(A) (its segment
  // tables: each listed row's entries in one panel are a segment): the
  // listed rows go segment by segment, B gathered from the panel's slice in
  // an XCD's L2, partial sums combined per row in panel order; the other
  // rows run the plain kernel.  Null: the plain kernel takes every row.
  const PanelPlan* panel = nullptr;
};
// Whether the plain SpMM kernel takes these operands (else the ordered ones).
bool spmm_plain_applies(const SpmmArgs& args);
// Workspace: the passes' split table (none for one pass), and with a panel
// plan the segments' partial sums and the combine's per-part row sums.
size_t spmm_workspace_bytes(const SpmmArgs& args);
hipError_t spmm_launch(const SpmmArgs& args, void* workspace, hipStream_t stream);

// ---- batched SpMV ----------------------------------------------------------
// Per-matrix table entry and work item of the persistent batched kernel
// (device-resident arrays built by the host plan).
struct BatchMatrix {
  const int64_t* offsets;
  const void* indices;
  const void* values;
  const void* x;
  void* y;
  int64_t incx;
  int64_t incy;
  double alpha;
  double beta;
};
struct BatchItem {
  int32_t matrix;
  int32_t lanes;  // lanes per row: 1..64, power of two; 0: CSR-stream through LDS
  int64_t row_begin;
  int64_t row_end;
};
// Entries a CSR-stream item stages in LDS (larger items take one lane per row).
constexpr int64_t kStreamEntries = 2048;
struct BatchLaunch {
  const BatchMatrix* mats;
  const BatchItem* items;
  const int64_t* ranges;  // work-group w walks items [ranges[w], ranges[w+1])
  int64_t nitems;
  int64_t grid;  // persistent workgroups
  Elem elem;     // F32 or F64
  int index_bytes;
};
hipError_t batched_spmv_launch(const BatchLaunch& launch, hipStream_t stream);

// ---- selector: features and MLP --------------------------------------------
size_t features_workspace_bytes();
hipError_t features_launch(const Compressed& a, float* out16_device, void* workspace,
                           hipStream_t stream);
size_t mlp_param_count(int k);
size_t mlp_scratch_floats(int64_t batch, int k);
hipError_t mlp_forward_launch(const float* params, int k, int64_t batch, const float* x,
                              float* y, float* scratch, hipStream_t stream);
hipError_t mlp_grad_launch(const float* params, int k, int64_t batch, const float* x,
                           const float* target, float* grads, float* scratch,
                           float* loss, hipStream_t stream);
hipError_t mlp_adam_launch(float* params, const float* grads, float* m, float* v,
                           int64_t n, float lr, float b1, float b2, float eps,
                           int64_t step, hipStream_t stream);
// y = (x - mean) / std over [n][16] rows, clamped to +-clip.
hipError_t normalize_launch(const float* x, const float* mean, const float* stddev,
                            int64_t n, float* out, hipStream_t stream);

// ---- conversions ---------------------------------------------------------
// Stable transpose of a compressed operand (CSR<->CSC): out_offsets has
// inner+1 entries, out_indices nnz entries of out_index_bytes, out_values
// nnz values of the same element kind.
size_t transpose_workspace_bytes(const Compressed& src);
hipError_t transpose_launch(const Compressed& src, int64_t* out_offsets,
                            void* out_indices, int out_index_bytes,
                            void* out_values, void* workspace,
                            hipStream_t stream);

// The two-pass transpose (transpose.hip) behind transpose_launch for inner
// extents of 2^9..2^22 (digits of at most 11 bits), nnz and outer below 2^32
// and at most 16 row segments; transpose2_applies says whether it runs.
bool transpose2_applies(const Compressed& src);
size_t transpose2_workspace_bytes(const Compressed& src);
hipError_t transpose2_launch(const Compressed& src, int64_t* out_offsets, void* out_indices,
                             int out_index_bytes, void* out_values, void* workspace,
                             hipStream_t stream);

// Canonical COO (interleaved int64 (row, col)) -> CSR (to_csc false) or CSC.
size_t coo_convert_workspace_bytes(int64_t rows, int64_t cols, int64_t nnz,
                                   bool to_csc, Elem elem);
hipError_t coo_convert_launch(const int64_t* coords, const void* values,
                              Elem elem, int64_t rows, int64_t cols,
                              int64_t nnz, bool to_csc, int64_t* out_offsets,
                              void* out_indices, int out_index_bytes,
                              void* out_values, void* workspace,
                              hipStream_t stream);

// Raw device triplets (interleaved int64 (row, col), F32/F64 values) ->
// canonical COO, as CoordinateBuilder::Finalize: stable (row, col) order,
// runs of equal coordinates summed in insertion order (or rejected), exact
// zeros dropped on request.  Synchronous: *out_nnz is the finalized count;
// *out_failure is 1 for a coordinate outside the shape, 2 for a duplicate
// under the reject policy (outputs untouched), else 0.  rows, cols and nnz
// must not exceed UINT32_MAX.
size_t coo_finalize_workspace_bytes(int64_t rows, int64_t cols, int64_t nnz, Elem elem);
hipError_t coo_finalize_launch(const int64_t* coords, const void* values, Elem elem,
                               int64_t rows, int64_t cols, int64_t nnz, bool sum_duplicates,
                               bool drop_zeros, int64_t* out_coords, void* out_values,
                               int64_t* out_nnz, int* out_failure, void* workspace,
                               hipStream_t stream);

// CSR (outer = rows) -> canonical COO: interleaved int64 (row, col) and values.
hipError_t csr_to_coo_launch(const Compressed& a, int64_t* coords, void* out_values,
                             hipStream_t stream);

// int64 -> int32 index narrowing (caller guarantees inner <= INT32_MAX).
hipError_t narrow_indices_launch(const int64_t* src, int32_t* dst, int64_t n,
                                 hipStream_t stream);
// FP64 -> FP32/BF16/F16 with round-to-nearest-even.
hipError_t convert_values_launch(const double* src, void* dst, Elem to,
                                 int64_t n, double scale, hipStream_t stream);
// Byte fill as a kernel (graph-capture safe, unlike hipMemsetAsync).
hipError_t fill_launch(void* dst, uint8_t byte, size_t bytes,
                       hipStream_t stream);

// Sparse Level-1 over a canonical indexed vector: op 0 dot (into `result`,
// one element; `ordered` = serial left-to-right), 1 axpy (x += alpha*v at the
// indices), 2 gather (v = x at the indices), 3 gather-zero (and zero x
// there), 4 scatter (x = v at the indices).  Dense element i is dense[i*inc].
hipError_t level1_launch(int op, bool ordered, Elem elem, double alpha, const int64_t* indices,
                         void* values, void* dense, int64_t inc, int64_t nnz, void* result,
                         hipStream_t stream);

size_t elem_bytes(Elem e);

}  // namespace asc_hip
