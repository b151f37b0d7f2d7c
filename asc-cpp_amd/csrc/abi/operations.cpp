// C ABI operations: staging, SpMV, SpMM and conversions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using asc_hip::Elem;
using namespace asc_hip_abi;

namespace {

constexpr int64_t kU32Max = std::numeric_limits<uint32_t>::max();

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int64_t outer_of(const asc_hip_sparse_desc& d) {
  return d.format == ASC_HIP_CSC ? d.cols : d.rows;
}
int64_t inner_of(const asc_hip_sparse_desc& d) {
  return d.format == ASC_HIP_CSC ? d.rows : d.cols;
}

template <typename I>
asc_hip_status validate_host_compressed_typed(const asc_hip_sparse_desc& d) {
  const int64_t outer = outer_of(d), inner = inner_of(d);
  const int64_t* off = static_cast<const int64_t*>(d.outer_offsets);
  const I* idx = static_cast<const I*>(d.inner_indices);
  // compressed.h:711-746
  if (off[0] != 0 || off[outer] != d.nnz)
    return fail(ASC_HIP_INVALID_ARGUMENT, "offsets must start at 0 and end at nnz");
  for (int64_t o = 0; o < outer; ++o) {
    const int64_t b = off[o], e = off[o + 1];
    if (e < b || e > d.nnz) return fail(ASC_HIP_INVALID_ARGUMENT, "offsets must be nondecreasing");
    for (int64_t p = b; p < e; ++p) {
      const int64_t v = static_cast<int64_t>(idx[p]);
      if (v < 0 || v >= inner) return fail(ASC_HIP_INDEX, "inner index out of range");
      if (p > b && v <= static_cast<int64_t>(idx[p - 1]))
        return fail(ASC_HIP_INVALID_ARGUMENT, "inner indices must increase strictly within a segment");
    }
  }
  return ok();
}

asc_hip_status validate_host_compressed(const asc_hip_sparse_desc& d) {
  if (d.index_bytes == 4) return validate_host_compressed_typed<int32_t>(d);
  return validate_host_compressed_typed<int64_t>(d);
}

bool host_space(uint8_t s) { return s == ASC_HIP_HOST || s == ASC_HIP_PINNED_HOST; }

// ---- SpMV preparation -----------------------------------------------------
struct SpmvPlan {
  asc_hip::SpmvArgs args{};
  bool transposed = false;  // op(A) needs a device transpose first
  asc_hip::Compressed source{};
  int64_t t_outer = 0;      // transposed operand's outer extent
  size_t t_offsets = 0, t_indices = 0, t_values = 0, t_ws = 0, spmv_ws = 0;
  size_t total = 0;
  bool noop = false;
  asc_hip::PanelPlan spmm_panel{};  // the plan an ASC_HIP_ALG_PANEL SpMM runs on
  bool spmm_panel_optional = false;  // AUTO took it: dropped when the workspace is short
};

// The algorithm a call runs: the explicit choice, else the plan cache's entry
// for the matrix, else merge-path.  `panel` receives the matrix's
// column-panel analysis when the cache holds one.
uint8_t resolve_alg(asc_hip_context* ctx, const asc_hip_sparse_desc& a,
                    const asc_hip_spmv_options* opt, asc_hip::PanelPlan* panel,
                    bool* has_panel) {
  const uint8_t req = opt ? opt->algorithm : ASC_HIP_ALG_AUTO;
  *has_panel = false;
  if (req != ASC_HIP_ALG_AUTO && req != ASC_HIP_ALG_PANEL) return req;
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  for (const auto& p : ctx->plans) {
    if (p.offsets == a.outer_offsets && p.indices == a.inner_indices &&
        p.rows == a.rows && p.cols == a.cols && p.nnz == a.nnz &&
        p.element_kind == a.element_kind) {
      if (p.algorithm == ASC_HIP_ALG_PANEL) {
        // The stream form runs on its copy of the values, which an in-place
        // change of the caller's values would leave stale: AUTO (the
        // reference's read-the-current-values semantics) never takes it and
        // runs merge-path; only an explicit ASC_HIP_ALG_PANEL does.
        if (req == ASC_HIP_ALG_AUTO && p.panel.stream) return ASC_HIP_ALG_MERGE;
        *panel = p.panel;
        *has_panel = true;
      }
      return req == ASC_HIP_ALG_AUTO ? p.algorithm : req;
    }
  }
  return req == ASC_HIP_ALG_AUTO ? ASC_HIP_ALG_MERGE : req;
}

// The plain SpMM gathers B one slice of kSpmmPassBytes at a time, one pass
// over the matrix per slice, so that the slice stays in the 256 MiB
// memory-side cache while its rows are gathered.  Measured on the 640M-entry
// Zipf matrix with a 1 GiB B (k = 32, FP64): one pass 27.2 ms, 512/256/128
// MiB slices 25.1/24.5/24.8 ms.  The ordered form on the same operands runs
// row by row with 16 lanes a row and takes the same passes (one pass
// 36.6 ms, 256 MiB slices 31.2 ms); the ordered kernels for other operands
// run in one pass (slices of 256/128/64 MiB made them 4%/11%/30% slower).
constexpr int64_t kSpmmPassBytes = int64_t(256) << 20;

void spmm_plain_plan(asc_hip::SpmmArgs* m, size_t esz) {
  m->panel_cols = 0;
  if (!asc_hip::spmm_plain_applies(*m) || m->k <= 0 || m->a.inner <= 0) return;
  m->panel_cols = std::max<int64_t>(1, kSpmmPassBytes / (m->k * static_cast<int64_t>(esz)));
}

asc_hip_status missing_panel_analysis() {
  return fail(ASC_HIP_INVALID_ARGUMENT,
              "ASC_HIP_ALG_PANEL needs a column-panel analysis of this matrix "
              "(asc_sparse_hip_csr_spmv_analyze)");
}

// Default panel: 2 MiB of x, half an XCD's L2 (the other half keeps the
// streamed matrix lines).  Rows this long or longer take the panel kernels.
constexpr int64_t kDefaultPanelBytes = int64_t(2) << 20;
constexpr int64_t kDefaultLongRowMin = 128;
constexpr int64_t kMaxPanels = 1024;
constexpr double kLdsMinSegment = 48.0;

struct PanelShape {
  int64_t panel_cols = 1;
  int64_t panels = 8;
  int64_t long_min = kDefaultLongRowMin;
  bool lds = false;
  bool stream = false;
  asc_hip::StreamLayout layout;  // stream staging only
  size_t list_bytes = 0, begin_bytes = 0, split_bytes = 0, scratch_bytes = 0, bytes = 0;
};

// Rows with at least long_min entries and their entries, counted on the host
// from the matrix's offsets (read back after the context's stream drains).
// A validation-only context has no device to read: every row counts.
asc_hip_status count_listed_rows(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                 int64_t long_min, int64_t panels, int64_t* nlong,
                                 int64_t* nnz_long, int64_t* segs_bound) {
  *nlong = a->rows;
  *nnz_long = a->nnz;
  *segs_bound = a->nnz;
  if (ctx->device < 0 || a->rows == 0) return ok();
  std::vector<int64_t> off(static_cast<size_t>(a->rows) + 1);
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  hipError_t e = hipMemcpyAsync(off.data(), a->outer_offsets, sizeof(int64_t) * off.size(),
                                hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail_hip(e, "reading the matrix offsets");
  // A row's entries in one panel form one segment: at most min(len, panels).
  int64_t n = 0, entries = 0, segs = 0;
  for (int64_t r = 0; r < a->rows; ++r) {
    const int64_t len = off[r + 1] - off[r];
    if (len >= long_min) {
      ++n;
      entries += len;
      segs += std::min(len, panels);
    }
  }
  *nlong = n;
  *nnz_long = entries;
  *segs_bound = segs;
  return ok();
}

asc_hip_status panel_shape(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                           const asc_hip_panel_options* opt, PanelShape* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format != ASC_HIP_CSR)
    return fail(ASC_HIP_INVALID_ARGUMENT, "column-panel analysis requires CSR format");
  if (a->element_kind == ASC_HIP_C64 || a->element_kind == ASC_HIP_C128)
    return fail(ASC_HIP_UNSUPPORTED, "compat SpMV is defined for real storage");
  const int64_t want_cols = opt ? opt->panel_columns : 0;
  const uint8_t staging_req = opt ? opt->staging : ASC_HIP_STAGING_AUTO;
  // Stream staging lists every nonempty row by default: its x comes from
  // LDS, so even short rows gain, and the combine walks rows in groups.
  const int64_t long_min = opt && opt->long_row_min ? opt->long_row_min
                           : staging_req == ASC_HIP_STAGING_STREAM ? 1
                                                                   : kDefaultLongRowMin;
  if (want_cols < 0 || long_min < 0)
    return fail(ASC_HIP_INVALID_ARGUMENT, "panel width and long-row length must be nonnegative");
  if (a->rows > kU32Max || std::min(a->nnz, a->cols) > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "column-panel analysis stores 32-bit row ids and positions");
  // The stream form records each segment's first entry as a 32-bit position.
  if (staging_req == ASC_HIP_STAGING_STREAM && a->nnz > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "the stream form records 32-bit entry positions");
  const uint8_t staging = opt ? opt->staging : ASC_HIP_STAGING_AUTO;
  if (staging > ASC_HIP_STAGING_STREAM)
    return fail(ASC_HIP_INVALID_ARGUMENT, "unknown panel staging");
  const int64_t xbytes = a->element_kind == ASC_HIP_F64 ? 8 : 4;
  if (staging == ASC_HIP_STAGING_STREAM) {
    // Panels of at most 128 KiB of x (the rest of the 160 KiB LDS holds the
    // waves' segment rings) and 32768 columns (15-bit column field).
    const int64_t max_cols = std::min<int64_t>(asc_hip::kStreamPanelBytes / xbytes,
                                               asc_hip::kStreamMaxPanelCols);
    if (want_cols > max_cols)
      return fail(ASC_HIP_INVALID_ARGUMENT,
                  "a stream panel holds at most 128 KiB of x and 32768 columns");
    const int64_t w = want_cols > 0 ? want_cols : max_cols;
    int64_t panels = a->cols > w ? (a->cols - 1) / w + 1 : 1;
    panels = (panels + 7) / 8 * 8;
    if (panels > kMaxPanels)
      return fail(ASC_HIP_INVALID_ARGUMENT, "panel width gives more than 1024 panels");
    int64_t nlong = 0, nnz_long = 0, segs = 0;
    if ((st = count_listed_rows(ctx, a, long_min, panels, &nlong, &nnz_long, &segs)).code)
      return st;
    if (nnz_long + panels * asc_hip::kStreamTile > static_cast<int64_t>(kU32Max))
      return fail(ASC_HIP_UNSUPPORTED, "the stream form numbers its entries in 32 bits");
    out->panels = panels;
    out->panel_cols = a->cols > 0 ? (a->cols - 1) / panels + 1 : 1;
    out->long_min = long_min;
    out->stream = true;
    out->layout = asc_hip::stream_layout(a->rows, nlong, nnz_long, segs, panels,
                                         elem_size(a->element_kind));
    out->bytes = out->layout.total;
    return ok();
  }
  // Panels of a multiple of 8 covering the columns with at most `w` each.
  auto cut = [&](int64_t w, int64_t* panels, int64_t* cols) {
    const int64_t p0 = a->cols > w ? (a->cols - 1) / w + 1 : 1;
    *panels = (p0 + 7) / 8 * 8;
    *cols = a->cols > 0 ? (a->cols - 1) / *panels + 1 : 1;
  };
  const int64_t lds_cols = asc_hip::kLdsPanelBytes / xbytes;
  int64_t panels = 0, panel_cols = 0;
  cut(want_cols > 0 ? want_cols : (staging == ASC_HIP_STAGING_LDS ? lds_cols
                                                                   : kDefaultPanelBytes / xbytes),
      &panels, &panel_cols);
  bool lds = staging == ASC_HIP_STAGING_LDS;
  if (lds && panel_cols > lds_cols)
    return fail(ASC_HIP_INVALID_ARGUMENT, "an LDS-staged panel holds at most 160 KiB of x");
  if (staging == ASC_HIP_STAGING_AUTO) {
    // LDS staging pays when a row's entries in one LDS-sized panel form
    // segments long enough to stream (measured crossover near 48 entries);
    // shorter segments stream better from fewer, L2-sized panels.
    int64_t lp = panels, lc = panel_cols;
    if (want_cols == 0) cut(lds_cols, &lp, &lc);
    const double per_segment = a->rows > 0 ? static_cast<double>(a->nnz) /
                                                 static_cast<double>(a->rows) /
                                                 static_cast<double>(lp)
                                           : 0.0;
    if (lc <= lds_cols && per_segment >= kLdsMinSegment) {
      lds = true;
      panels = lp;
      panel_cols = lc;
    }
  }
  if (panels > kMaxPanels)
    return fail(ASC_HIP_INVALID_ARGUMENT, "panel width gives more than 1024 panels");
  int64_t split_entries = 0;
  if (!checked_mul(panels, a->rows, &split_entries) ||
      split_entries > std::numeric_limits<int64_t>::max() / 8)
    return fail(ASC_HIP_OVERFLOW, "panel split table size overflows");
  out->panels = panels;
  out->panel_cols = panel_cols;
  out->lds = lds;
  out->long_min = long_min;
  out->list_bytes = align256(sizeof(uint32_t) * static_cast<size_t>(a->rows));
  out->begin_bytes = align256(sizeof(int64_t) * static_cast<size_t>(a->rows));
  out->split_bytes = align256(sizeof(uint32_t) * static_cast<size_t>(split_entries));
  out->scratch_bytes = asc_hip::panel_analysis_scratch_bytes(a->rows);
  out->bytes = out->list_bytes + out->begin_bytes + out->split_bytes + out->scratch_bytes;
  return ok();
}

// The row list and split table must survive the SpMV's own writes.
bool panel_overlaps(const asc_hip::PanelPlan& pp, Span out) {
  if (pp.stream) return overlaps(span_of(pp.stream_base, pp.stream_bytes), out);
  if (pp.nlong == 0) return false;
  const size_t n = static_cast<size_t>(pp.nlong);
  return overlaps(span_of(pp.long_rows, 4 * n), out) ||
         overlaps(span_of(pp.row_begin, 8 * n), out) ||
         overlaps(span_of(pp.splits, 4 * n * static_cast<size_t>(pp.panels)), out);
}

// The stream form runs on its copy of the analysed (or last refreshed)
// values, with plain accumulation.
asc_hip_status check_stream_plan(const asc_hip_sparse_desc& a, const asc_hip::PanelPlan& pp,
                                 const asc_hip_spmv_options* opt) {
  if (!pp.stream) return ok();
  if (a.values != pp.values_src)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "the stream-staged panel analysis holds a copy of other values; refresh it "
                "(asc_sparse_hip_csr_spmv_refresh_values) or analyze the matrix again");
  if (opt && opt->accumulate == ASC_HIP_ACC_TWOSUM)
    return fail(ASC_HIP_UNSUPPORTED, "the stream-staged panel plan accumulates plainly");
  return ok();
}

asc_hip_status check_options(const asc_hip_sparse_desc& a, const asc_hip_spmv_options* opt,
                             uint8_t alg) {
  if (alg == ASC_HIP_ALG_AUTO || alg >= ASC_HIP_ALG_COUNT)
    return fail(ASC_HIP_INVALID_ARGUMENT, "unknown SpMV algorithm");
  const uint8_t acc = opt ? opt->accumulate : ASC_HIP_ACC_PLAIN;
  if (acc > ASC_HIP_ACC_TWOSUM) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown accumulation mode");
  if (acc == ASC_HIP_ACC_TWOSUM && a.element_kind == ASC_HIP_F64)
    return fail(ASC_HIP_UNSUPPORTED, "compensated accumulation applies to FP32-accumulated storage");
  if (alg == ASC_HIP_ALG_ORDERED &&
      (acc != ASC_HIP_ACC_PLAIN ||
       (a.element_kind != ASC_HIP_F32 && a.element_kind != ASC_HIP_F64)))
    return fail(ASC_HIP_UNSUPPORTED, "ordered summation is defined for FP32/FP64 storage");
  return ok();
}

asc_hip_status prepare_compat_spmv(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                   const asc_hip_vector_desc* x,
                                   const asc_hip_vector_desc* y,
                                   const asc_hip_spmv_options* opt, SpmvPlan* plan) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format != ASC_HIP_CSR) return fail(ASC_HIP_INVALID_ARGUMENT, "CSR SpMV requires CSR format");
  if (a->element_kind == ASC_HIP_C64 || a->element_kind == ASC_HIP_C128)
    return fail(ASC_HIP_UNSUPPORTED, "compat SpMV is defined for real storage");
  const uint8_t vk = a->element_kind == ASC_HIP_F64 ? ASC_HIP_F64 : ASC_HIP_F32;
  if ((st = check_vector(ctx, x, vk, a->cols, false, "input")).code) return st;
  if ((st = check_vector(ctx, y, vk, a->rows, true, "output")).code) return st;
  const Span ys = vector_span(*y);
  if (overlaps(vector_span(*x), ys) || sparse_overlaps(*a, ys))
    return fail(ASC_HIP_INVALID_ARGUMENT, "SpMV rejects output operand overlap");
  asc_hip::PanelPlan panel;
  bool has_panel = false;
  const uint8_t alg = resolve_alg(ctx, *a, opt, &panel, &has_panel);
  if ((st = check_options(*a, opt, alg)).code) return st;
  if (alg == ASC_HIP_ALG_PANEL && !has_panel) return missing_panel_analysis();
  if (alg == ASC_HIP_ALG_PANEL && panel_overlaps(panel, ys))
    return fail(ASC_HIP_INVALID_ARGUMENT, "SpMV output overlaps the column-panel analysis");
  if (alg == ASC_HIP_ALG_PANEL && (st = check_stream_plan(*a, panel, opt)).code) return st;
  asc_hip::SpmvArgs& g = plan->args;
  g.panel = panel;
  g.a = as_compressed(*a, false);
  g.x = x->data;
  g.incx = x->stride;
  g.y = const_cast<void*>(y->data);
  g.incy = y->stride;
  g.form = asc_hip::SpmvForm::Compat;
  g.alg = static_cast<asc_hip::SpmvAlg>(alg);
  g.twosum = opt && opt->accumulate == ASC_HIP_ACC_TWOSUM;
  plan->spmv_ws = asc_hip::spmv_workspace_bytes(g);
  plan->total = plan->spmv_ws;
  plan->noop = a->rows == 0;
  return ok();
}

// Standardized SpMV/SpMM: the kernels consume op(A) as a row view.  CSR/N
// and CSC/T,H are row views already; CSR/T,H and CSC/N are transposed into
// the workspace first.
asc_hip_status plan_row_view(const asc_hip_sparse_desc& a, uint8_t transpose,
                             SpmvPlan* plan) {
  const bool csr = a.format == ASC_HIP_CSR;
  const bool direct = (csr && transpose == ASC_HIP_OP_N) || (!csr && transpose != ASC_HIP_OP_N);
  plan->source = as_compressed(a, false);
  if (direct) {
    plan->args.a = plan->source;
    return ok();
  }
  const int64_t n = a.nnz, outer = outer_of(a), inner = inner_of(a);
  if (n > kU32Max || outer > kU32Max || inner > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "device transpose supports extents and nnz below 2^32");
  plan->transposed = true;
  plan->t_outer = inner;
  const size_t esz = elem_size(a.element_kind);
  plan->t_offsets = align256(sizeof(int64_t) * (inner + 1));
  plan->t_indices = align256(static_cast<size_t>(a.index_bytes) * n);
  plan->t_values = align256(esz * n);
  plan->t_ws = align256(asc_hip::transpose_workspace_bytes(plan->source));
  asc_hip::Compressed rv;
  rv.offsets = nullptr;
  rv.indices = nullptr;
  rv.values = nullptr;
  rv.outer = inner;
  rv.inner = outer;
  rv.nnz = n;
  rv.index_bytes = a.index_bytes;
  rv.elem = plan->source.elem;
  plan->args.a = rv;
  return ok();
}

// Points the row view at the transposed arrays inside the workspace and
// enqueues the transpose.
hipError_t run_transpose(SpmvPlan* plan, char* ws, hipStream_t s) {
  int64_t* off = reinterpret_cast<int64_t*>(ws);
  void* idx = ws + plan->t_offsets;
  void* val = ws + plan->t_offsets + plan->t_indices;
  void* tws = ws + plan->t_offsets + plan->t_indices + plan->t_values;
  plan->args.a.offsets = off;
  plan->args.a.indices = idx;
  plan->args.a.values = val;
  return asc_hip::transpose_launch(plan->source, off, idx, plan->source.index_bytes,
                                   val, tws, s);
}

size_t transpose_prefix(const SpmvPlan& p) {
  return p.transposed ? p.t_offsets + p.t_indices + p.t_values + p.t_ws : 0;
}

asc_hip_status prepare_standard_spmv(asc_hip_context* ctx, uint8_t transpose,
                                     const asc_hip_sparse_desc* a,
                                     const asc_hip_vector_desc* x,
                                     const asc_hip_vector_desc* y,
                                     const asc_hip_spmv_options* opt, SpmvPlan* plan) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (transpose > ASC_HIP_OP_H) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown transpose");
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "standard SpMV requires CSR or CSC");
  if (a->element_kind != ASC_HIP_F32 && a->element_kind != ASC_HIP_F64)
    return fail(ASC_HIP_UNSUPPORTED, "standard SpMV supports real FP32/FP64 storage");
  const int64_t in_ext = transpose == ASC_HIP_OP_N ? a->cols : a->rows;
  const int64_t out_ext = transpose == ASC_HIP_OP_N ? a->rows : a->cols;
  if ((st = check_vector(ctx, x, a->element_kind, in_ext, false, "input")).code) return st;
  if ((st = check_vector(ctx, y, a->element_kind, out_ext, true, "output")).code) return st;
  const Span ys = vector_span(*y);
  if (overlaps(vector_span(*x), ys) || sparse_overlaps(*a, ys))
    return fail(ASC_HIP_INVALID_ARGUMENT, "Sparse matrix-vector input and output overlap");
  asc_hip::PanelPlan panel;
  bool has_panel = false;
  uint8_t alg = resolve_alg(ctx, *a, opt, &panel, &has_panel);
  if ((st = check_options(*a, opt, alg)).code) return st;
  if (alg == ASC_HIP_ALG_PANEL) {
    if (!has_panel) return missing_panel_analysis();
    // The analysis describes the stored CSR rows: other row views run
    // merge-path.
    if (a->format != ASC_HIP_CSR || transpose != ASC_HIP_OP_N)
      alg = ASC_HIP_ALG_MERGE;
    else if (panel_overlaps(panel, ys))
      return fail(ASC_HIP_INVALID_ARGUMENT, "SpMV output overlaps the column-panel analysis");
    else if ((st = check_stream_plan(*a, panel, opt)).code)
      return st;
  }
  if ((st = plan_row_view(*a, transpose, plan)).code) return st;
  asc_hip::SpmvArgs& g = plan->args;
  g.panel = panel;
  g.x = x->data;
  g.incx = x->stride;
  g.y = const_cast<void*>(y->data);
  g.incy = y->stride;
  g.alpha = 0.0;
  g.beta = 0.0;
  g.form = asc_hip::SpmvForm::Standard;
  g.alg = static_cast<asc_hip::SpmvAlg>(alg);
  g.twosum = false;
  plan->spmv_ws = align256(asc_hip::spmv_workspace_bytes(g));
  plan->total = transpose_prefix(*plan) + plan->spmv_ws;
  plan->noop = out_ext == 0;
  return ok();
}

SpanList operand_spans(const asc_hip_sparse_desc& a, std::initializer_list<Span> more) {
  const SparseSpans s = sparse_spans(a);
  SpanList v{s.first, s.second, s.values};
  for (const Span& m : more) v.push(m);
  return v;
}

// AUTO lanes for a panel plan's short rows: the power of two at or below
// long_min / 8, between 1 and 64.
uint8_t auto_short_alg(int64_t long_min) {
  uint8_t alg = ASC_HIP_ALG_VECTOR1;
  for (int64_t g = 2; g <= long_min / 8 && alg < ASC_HIP_ALG_VECTOR64; g *= 2) ++alg;
  return alg;
}

// Maps the matrix to ASC_HIP_ALG_PANEL with `pp` in the context's plan cache.
asc_hip_status record_panel_plan(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                 const asc_hip::PanelPlan& pp) {
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  for (auto& p : ctx->plans) {
    if (p.offsets == a->outer_offsets && p.indices == a->inner_indices && p.rows == a->rows &&
        p.cols == a->cols && p.nnz == a->nnz && p.element_kind == a->element_kind) {
      p.algorithm = ASC_HIP_ALG_PANEL;
      p.panel = pp;
      return ok();
    }
  }
  ctx->plans.push_back({a->outer_offsets, a->inner_indices, a->rows, a->cols, a->nnz,
                        a->element_kind, ASC_HIP_ALG_PANEL, pp});
  return ok();
}

asc_hip_status run_spmv(asc_hip_context* ctx, SpmvPlan& plan, void* ws) {
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  char* w = static_cast<char*>(ws);
  if (plan.transposed) {
    const hipError_t e = run_transpose(&plan, w, ctx->stream);
    if (e != hipSuccess) return finish_launch(ctx, e, "transpose");
  }
  plan.args.fault = ctx->fault_word;
  if (plan.args.alg == asc_hip::SpmvAlg::Panel && plan.args.panel.stream &&
      test_hook_consume_stream_stall())
    plan.args.spin_limit = 0;
  const hipError_t e =
      asc_hip::spmv_launch(plan.args, w + transpose_prefix(plan), ctx->stream, &ctx->side);
  return finish_launch(ctx, e, "SpMV launch");
}

}  // namespace


extern "C" {

// ---- staging ---------------------------------------------------------------

asc_hip_status asc_sparse_hip_clone_compressed(asc_hip_context* ctx,
                                               const asc_hip_sparse_desc* src,
                                               uint8_t index_bytes, void* d_off,
                                               void* d_idx, void* d_val,
                                               asc_hip_sparse_desc* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  if ((st = check_sparse_meta(src, "source")).code) return st;
  if (src->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "use asc_sparse_hip_clone_coordinate for COO");
  if (!host_space(src->memory_space)) return fail(ASC_HIP_MEMORY_ACCESS, "clone source must be host memory");
  if (index_bytes != 4 && index_bytes != 8) return fail(ASC_HIP_INVALID_ARGUMENT, "index width must be 4 or 8 bytes");
  const int64_t outer = outer_of(*src), inner = inner_of(*src), n = src->nnz;
  if (index_bytes == 4 && inner > std::numeric_limits<int32_t>::max())
    return fail(ASC_HIP_OVERFLOW, "narrowing clone: inner extent exceeds INT32_MAX");
  if ((st = validate_host_compressed(*src)).code) return st;
  const size_t esz = elem_size(src->element_kind);
  const size_t off_b = sizeof(int64_t) * (outer + 1), idx_b = index_bytes * n, val_b = esz * n;
  if ((st = check_device_pointer(ctx, d_off, "offsets destination", off_b)).code) return st;
  if (n > 0) {
    if ((st = check_device_pointer(ctx, d_idx, "indices destination", idx_b)).code) return st;
    if ((st = check_device_pointer(ctx, d_val, "values destination", val_b)).code) return st;
  }
  if (overlaps(span_of(d_off, off_b), span_of(d_idx, idx_b)) ||
      overlaps(span_of(d_off, off_b), span_of(d_val, val_b)) ||
      overlaps(span_of(d_idx, idx_b), span_of(d_val, val_b)))
    return fail(ASC_HIP_INVALID_ARGUMENT, "clone destinations overlap");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e = hipMemcpyAsync(d_off, src->outer_offsets, off_b, hipMemcpyHostToDevice, ctx->stream);
  std::vector<int32_t> narrow32;
  std::vector<int64_t> widen64;
  if (e == hipSuccess && n > 0) {
    const void* idx_src = src->inner_indices;
    if (index_bytes != src->index_bytes) {
      if (index_bytes == 4) {
        narrow32.resize(n);
        const int64_t* s64 = static_cast<const int64_t*>(src->inner_indices);
        for (int64_t i = 0; i < n; ++i) narrow32[i] = static_cast<int32_t>(s64[i]);
        idx_src = narrow32.data();
      } else {
        widen64.resize(n);
        const int32_t* s32 = static_cast<const int32_t*>(src->inner_indices);
        for (int64_t i = 0; i < n; ++i) widen64[i] = s32[i];
        idx_src = widen64.data();
      }
    }
    e = hipMemcpyAsync(d_idx, idx_src, idx_b, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_val, src->values, val_b, hipMemcpyHostToDevice, ctx->stream);
  }
  // Host staging buffers die with this call: complete the copies first.
  const hipError_t sync = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = sync;
  if (e != hipSuccess) return finish_launch(ctx, e, "clone copy");
  *out = *src;
  out->outer_offsets = d_off;
  out->inner_indices = d_idx;
  out->values = d_val;
  out->index_bytes = index_bytes;
  out->offset_bytes = 8;
  out->memory_space = ASC_HIP_DEVICE;
  out->trusted = 1;
  return ok();
}

asc_hip_status asc_sparse_hip_clone_coordinate(asc_hip_context* ctx,
                                               const asc_hip_sparse_desc* src,
                                               void* d_coords, void* d_val,
                                               asc_hip_sparse_desc* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  if ((st = check_sparse_meta(src, "source")).code) return st;
  if (src->format != ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "coordinate clone requires COO");
  if (!host_space(src->memory_space)) return fail(ASC_HIP_MEMORY_ACCESS, "clone source must be host memory");
  // coordinate.h:751-782: in bounds and strictly lexicographic.
  const int64_t* c = static_cast<const int64_t*>(src->outer_offsets);
  for (int64_t p = 0; p < src->nnz; ++p) {
    const int64_t r = c[2 * p], col = c[2 * p + 1];
    if (r < 0 || r >= src->rows || col < 0 || col >= src->cols)
      return fail(ASC_HIP_INDEX, "coordinate out of range");
    if (p > 0) {
      const int64_t pr = c[2 * p - 2], pc = c[2 * p - 1];
      if (r < pr || (r == pr && col <= pc))
        return fail(ASC_HIP_INVALID_ARGUMENT, "coordinates must be canonical (sorted, unique)");
    }
  }
  const size_t cb = 16 * static_cast<size_t>(src->nnz), vb = elem_size(src->element_kind) * src->nnz;
  if (src->nnz > 0) {
    if ((st = check_device_pointer(ctx, d_coords, "coordinates destination", cb)).code) return st;
    if ((st = check_device_pointer(ctx, d_val, "values destination", vb)).code) return st;
    if (overlaps(span_of(d_coords, cb), span_of(d_val, vb)))
      return fail(ASC_HIP_INVALID_ARGUMENT, "clone destinations overlap");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipError_t e = hipSuccess;
  if (src->nnz > 0) {
    e = hipMemcpyAsync(d_coords, c, cb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_val, src->values, vb, hipMemcpyHostToDevice, ctx->stream);
  }
  const hipError_t sync = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = sync;
  if (e != hipSuccess) return finish_launch(ctx, e, "clone copy");
  *out = *src;
  out->outer_offsets = d_coords;
  out->inner_indices = nullptr;
  out->values = d_val;
  out->memory_space = ASC_HIP_DEVICE;
  out->trusted = 1;
  return ok();
}

// compressed.h:472-516, coordinate.h:533-569: the new values must form a
// non-wrapping span, must not overlap the structure, and may only overlap the
// current values by being exactly them.
asc_hip_status asc_sparse_hip_rebind_values(const asc_hip_sparse_desc* source,
                                            const void* values, asc_hip_sparse_desc* out) {
  asc_hip_status st = check_sparse_meta(source, "rebound matrix");
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  const size_t esz = elem_size(source->element_kind);
  if (source->nnz > 0) {
    if (static_cast<uint64_t>(source->nnz) > SIZE_MAX / esz)
      return fail(ASC_HIP_OVERFLOW, "value byte count overflows");
    const size_t bytes = static_cast<size_t>(source->nnz) * esz;
    if (values == nullptr)
      return fail(ASC_HIP_INVALID_ARGUMENT, "Nonempty sparse storage cannot be null");
    if (bytes > UINTPTR_MAX - reinterpret_cast<uintptr_t>(values))
      return fail(ASC_HIP_OVERFLOW, "Sparse storage address range overflows uintptr_t");
    const SparseSpans s = sparse_spans(*source);
    const Span v = span_of(values, bytes);
    if (overlaps(s.first, v) || overlaps(s.second, v))
      return fail(ASC_HIP_INVALID_ARGUMENT,
                  source->format == ASC_HIP_COO
                      ? "Rebound sparse values cannot overlap coordinates"
                      : "Rebound sparse values cannot overlap compressed structure");
    if (values != source->values && overlaps(s.values, v))
      return fail(ASC_HIP_INVALID_ARGUMENT,
                  "Rebound sparse values cannot partially overlap existing values");
  }
  asc_hip_sparse_desc r = *source;
  r.values = values;
  *out = r;
  return ok();
}

asc_hip_status asc_sparse_hip_narrow_indices(asc_hip_context* ctx,
                                             const asc_hip_sparse_desc* src,
                                             void* d_idx32, asc_hip_sparse_desc* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  if ((st = check_device_sparse(ctx, src, "source")).code) return st;
  if (src->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "narrowing applies to CSR/CSC");
  if (src->index_bytes != 8) return fail(ASC_HIP_INVALID_ARGUMENT, "source indices are not int64");
  if (inner_of(*src) > std::numeric_limits<int32_t>::max())
    return fail(ASC_HIP_OVERFLOW, "inner extent exceeds INT32_MAX");
  if (src->nnz > 0) {
    if ((st = check_device_pointer(ctx, d_idx32, "indices destination",
                                   sizeof(int32_t) * static_cast<size_t>(src->nnz))).code)
      return st;
    if (sparse_overlaps(*src, span_of(d_idx32, 4 * static_cast<size_t>(src->nnz))))
      return fail(ASC_HIP_INVALID_ARGUMENT, "destination overlaps the source");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  st = finish_launch(ctx, asc_hip::narrow_indices_launch(
      static_cast<const int64_t*>(src->inner_indices), static_cast<int32_t*>(d_idx32),
      src->nnz, ctx->stream), "narrow indices");
  if (st.code) return st;
  *out = *src;
  out->inner_indices = d_idx32;
  out->index_bytes = 4;
  return ok();
}

asc_hip_status asc_sparse_hip_convert_values(asc_hip_context* ctx,
                                             const asc_hip_sparse_desc* src,
                                             uint8_t kind, void* d_val,
                                             asc_hip_sparse_desc* out) {
  return asc_sparse_hip_convert_values_scaled(ctx, src, kind, 1.0, d_val, out);
}

asc_hip_status asc_sparse_hip_convert_values_scaled(asc_hip_context* ctx,
                                                    const asc_hip_sparse_desc* src,
                                                    uint8_t kind, double scale,
                                                    void* d_val,
                                                    asc_hip_sparse_desc* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  if ((st = check_device_sparse(ctx, src, "source")).code) return st;
  if (src->element_kind != ASC_HIP_F64) return fail(ASC_HIP_UNSUPPORTED, "value conversion reads FP64 storage");
  if (kind != ASC_HIP_F32 && kind != ASC_HIP_BF16 && kind != ASC_HIP_F16)
    return fail(ASC_HIP_INVALID_ARGUMENT, "target must be FP32, BF16 or F16");
  if (!(scale > 0.0) || scale > 1e300)
    return fail(ASC_HIP_INVALID_ARGUMENT, "scale must be positive and finite");
  if (src->nnz > 0) {
    if ((st = check_device_pointer(ctx, d_val, "values destination",
                                   elem_size(kind) * static_cast<size_t>(src->nnz))).code)
      return st;
    if (sparse_overlaps(*src, span_of(d_val, elem_size(kind) * src->nnz)))
      return fail(ASC_HIP_INVALID_ARGUMENT, "destination overlaps the source");
  }
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  st = finish_launch(ctx, asc_hip::convert_values_launch(
      static_cast<const double*>(src->values), d_val, static_cast<Elem>(kind), src->nnz,
      scale, ctx->stream), "convert values");
  if (st.code) return st;
  *out = *src;
  out->values = d_val;
  out->element_kind = kind;
  return ok();
}

// ---- SpMV --------------------------------------------------------------------

asc_hip_status asc_sparse_hip_csr_spmv_workspace_size(
    asc_hip_context* ctx, const asc_hip_sparse_desc* a, const asc_hip_vector_desc* x,
    const asc_hip_vector_desc* y, const asc_hip_spmv_options* opt, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  SpmvPlan plan;
  asc_hip_status st = prepare_compat_spmv(ctx, a, x, y, opt, &plan);
  if (st.code) return st;
  *out = plan.total;
  return ok();
}

asc_hip_status asc_sparse_hip_csr_spmv(asc_hip_context* ctx, double alpha,
                                       const asc_hip_sparse_desc* a,
                                       const asc_hip_vector_desc* x, double beta,
                                       const asc_hip_vector_desc* y, void* ws,
                                       size_t ws_bytes, const asc_hip_spmv_options* opt) {
  SpmvPlan plan;
  asc_hip_status st = prepare_compat_spmv(ctx, a, x, y, opt, &plan);
  if (st.code) return st;
  if ((st = check_workspace(ctx, ws, ws_bytes, plan.total,
                            operand_spans(*a, {vector_span(*x), vector_span(*y)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (plan.noop) return ok();
  plan.args.alpha = alpha;
  plan.args.beta = beta;
  return run_spmv(ctx, plan, ws);
}

// ---- allocation-free compat SpMV ------------------------------------------

asc_hip_status asc_sparse_hip_csr_spmv_noalloc(asc_hip_context* ctx, double alpha,
                                               const asc_hip_sparse_desc* a,
                                               const asc_hip_vector_desc* x, double beta,
                                               const asc_hip_vector_desc* y,
                                               const asc_hip_spmv_options* opt) {
  SpmvPlan plan;
  asc_hip_status st = prepare_compat_spmv(ctx, a, x, y, opt, &plan);
  if (st.code) return st;
  if (plan.total > ASC_HIP_NOALLOC_SCRATCH_BYTES)
    return fail(ASC_HIP_UNSUPPORTED,
                "this SpMV needs " + std::to_string(plan.total) +
                    " bytes of workspace, beyond the allocation-free scratch (" +
                    std::to_string(ASC_HIP_NOALLOC_SCRATCH_BYTES) +
                    " bytes): use asc_sparse_hip_csr_spmv with a workspace");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (plan.noop) return ok();
  if (ctx->noalloc_scratch == nullptr)
    return fail(ASC_HIP_INVALID_STATE, "the allocation-free scratch is not set up");
  plan.args.alpha = alpha;
  plan.args.beta = beta;
  return run_spmv(ctx, plan, ctx->noalloc_scratch);
}

// ---- column-panel analysis -----------------------------------------------

asc_hip_status asc_sparse_hip_csr_spmv_analysis_size(asc_hip_context* ctx,
                                                     const asc_hip_sparse_desc* a,
                                                     const asc_hip_panel_options* opt,
                                                     size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  PanelShape shape;
  asc_hip_status st = panel_shape(ctx, a, opt, &shape);
  if (st.code) return st;
  *out = shape.bytes;
  return ok();
}

asc_hip_status asc_sparse_hip_csr_spmv_analyze(asc_hip_context* ctx,
                                               const asc_hip_sparse_desc* a,
                                               const asc_hip_panel_options* opt,
                                               void* analysis, size_t analysis_bytes) {
  PanelShape shape;
  asc_hip_status st = panel_shape(ctx, a, opt, &shape);
  if (st.code) return st;
  uint8_t short_alg = opt ? opt->short_algorithm : ASC_HIP_ALG_AUTO;
  uint8_t long_alg = opt ? opt->long_algorithm : ASC_HIP_ALG_AUTO;
  for (uint8_t alg : {short_alg, long_alg})
    if (alg != ASC_HIP_ALG_AUTO && (alg < ASC_HIP_ALG_VECTOR1 || alg > ASC_HIP_ALG_VECTOR64))
      return fail(ASC_HIP_INVALID_ARGUMENT, "panel plans run ASC_HIP_ALG_VECTOR1..VECTOR64");
  const SparseSpans sp = sparse_spans(*a);
  if ((st = check_workspace(ctx, analysis, analysis_bytes, shape.bytes,
                            {sp.first, sp.second, sp.values})).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (shape.stream) {
    asc_hip::PanelPlan pp;
    pp.panels = shape.panels;
    pp.panel_cols = shape.panel_cols;
    pp.long_min = shape.long_min;
    {
      DeviceGuard guard(ctx->device);
      if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
      const hipError_t e = asc_hip::panel_stream_analysis_launch(
          as_compressed(*a, false), shape.layout, analysis, &pp, ctx->stream);
      if ((st = finish_launch(ctx, e, "stream panel analysis")).code) return st;
    }
    if (short_alg == ASC_HIP_ALG_AUTO) short_alg = auto_short_alg(shape.long_min);
    pp.short_alg = static_cast<asc_hip::SpmvAlg>(short_alg);
    pp.long_alg = asc_hip::SpmvAlg::Vec64;  // unused by the stream kernels
    return record_panel_plan(ctx, a, pp);
  }
  char* w = static_cast<char*>(analysis);
  uint32_t* long_rows = reinterpret_cast<uint32_t*>(w);
  int64_t* row_begin = reinterpret_cast<int64_t*>(w + shape.list_bytes);
  uint32_t* splits = reinterpret_cast<uint32_t*>(w + shape.list_bytes + shape.begin_bytes);
  void* scratch = w + shape.list_bytes + shape.begin_bytes + shape.split_bytes;
  asc_hip::PanelPlan pp;
  pp.long_rows = long_rows;
  pp.row_begin = row_begin;
  pp.splits = splits;
  pp.panels = shape.panels;
  pp.panel_cols = shape.panel_cols;
  pp.long_min = shape.long_min;
  pp.lds = shape.lds;
  {
    DeviceGuard guard(ctx->device);
    if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
    int64_t nlong = 0;
    const hipError_t e = asc_hip::panel_analysis_launch(
        as_compressed(*a, false), shape.panel_cols, shape.panels, shape.long_min, long_rows,
        row_begin, splits, scratch, &nlong, ctx->stream);
    if ((st = finish_launch(ctx, e, "panel analysis")).code) return st;
    pp.nlong = nlong;
  }
  // LDS form: row chunks so that each work-group gathers about eight entries
  // per staged x element (the panel copy stays a small share of its reads),
  // with at least one work-group per CU.
  if (pp.lds && pp.nlong > 0) {
    const int64_t per_chunk = std::max<int64_t>(1, 8 * a->cols);
    int64_t chunks = std::max<int64_t>(1, a->nnz / per_chunk);
    chunks = std::max<int64_t>(chunks, (256 + pp.panels - 1) / pp.panels);
    chunks = std::min<int64_t>(chunks, pp.nlong);
    chunks = std::min<int64_t>(chunks, std::numeric_limits<uint32_t>::max() / pp.panels);
    pp.chunks = chunks;
  }
  // AUTO: for the short rows, the power of two at or below long_min / 8
  // lanes; for the listed rows, the power of two at or below their mean
  // entries per panel (every entry counted as a long row's: most are, in the
  // skewed matrices this plan is for).  Both between 1 and 64.
  if (short_alg == ASC_HIP_ALG_AUTO) short_alg = auto_short_alg(shape.long_min);
  if (long_alg == ASC_HIP_ALG_AUTO) {
    // LDS staging: about six entries per lane (a pass's metadata loads are
    // then a small share of its work); L2 staging: one entry per lane.
    const double mean = pp.nlong > 0 ? static_cast<double>(a->nnz) /
                                           static_cast<double>(pp.nlong) /
                                           static_cast<double>(pp.panels) /
                                           (pp.lds ? 6.0 : 1.0)
                                     : 1.0;
    long_alg = ASC_HIP_ALG_VECTOR1;
    for (double g = 2.0; g <= mean && long_alg < ASC_HIP_ALG_VECTOR64; g *= 2.0) ++long_alg;
  }
  pp.short_alg = static_cast<asc_hip::SpmvAlg>(short_alg);
  pp.long_alg = static_cast<asc_hip::SpmvAlg>(long_alg);
  return record_panel_plan(ctx, a, pp);
}

asc_hip_status asc_sparse_hip_csr_spmv_refresh_values(asc_hip_context* ctx,
                                                      const asc_hip_sparse_desc* a) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format != ASC_HIP_CSR)
    return fail(ASC_HIP_INVALID_ARGUMENT, "column-panel plans are built for CSR matrices");
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  for (auto& p : ctx->plans) {
    if (p.offsets == a->outer_offsets && p.indices == a->inner_indices && p.rows == a->rows &&
        p.cols == a->cols && p.nnz == a->nnz && p.element_kind == a->element_kind) {
      if (p.algorithm != ASC_HIP_ALG_PANEL || !p.panel.stream)
        return fail(ASC_HIP_INVALID_STATE,
                    "the matrix has no stream-staged panel analysis in this context");
      if (overlaps(span_of(p.panel.stream_base, p.panel.stream_bytes), sparse_spans(*a).values))
        return fail(ASC_HIP_INVALID_ARGUMENT, "the values overlap the panel analysis");
      DeviceGuard guard(ctx->device);
      if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
      const hipError_t e = asc_hip::panel_stream_refresh_launch(
          p.panel, a->values, elem_size(a->element_kind), ctx->stream);
      if ((st = finish_launch(ctx, e, "stream panel value refresh")).code) return st;
      p.panel.values_src = a->values;
      return ok();
    }
  }
  return fail(ASC_HIP_INVALID_STATE, "the matrix has no stream-staged panel analysis in this context");
}

asc_hip_status asc_sparse_hip_standard_spmv_workspace_size(
    asc_hip_context* ctx, uint8_t transpose, const asc_hip_sparse_desc* a,
    const asc_hip_vector_desc* x, const asc_hip_vector_desc* y,
    const asc_hip_spmv_options* opt, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  SpmvPlan plan;
  asc_hip_status st = prepare_standard_spmv(ctx, transpose, a, x, y, opt, &plan);
  if (st.code) return st;
  *out = plan.total;
  return ok();
}

asc_hip_status asc_sparse_hip_standard_spmv(asc_hip_context* ctx, uint8_t transpose,
                                            asc_hip_scalar alpha,
                                            const asc_hip_sparse_desc* a,
                                            const asc_hip_vector_desc* x,
                                            const asc_hip_vector_desc* y, void* ws,
                                            size_t ws_bytes,
                                            const asc_hip_spmv_options* opt) {
  SpmvPlan plan;
  asc_hip_status st = prepare_standard_spmv(ctx, transpose, a, x, y, opt, &plan);
  if (st.code) return st;
  if ((st = check_workspace(ctx, ws, ws_bytes, plan.total,
                            operand_spans(*a, {vector_span(*x), vector_span(*y)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  // blas.h:1650-1652: alpha == 0 returns without reading anything.
  if (alpha.real == 0.0 || plan.noop) return ok();
  plan.args.alpha = alpha.real;
  return run_spmv(ctx, plan, ws);
}

// ---- SpMM --------------------------------------------------------------------

namespace {

// The SpMM operands of a call, in kernel form (B and C pointers included, so
// that the plain kernel's alignment conditions are known at sizing time).
void spmm_args(const SpmvPlan& plan, const asc_hip_matrix_desc& b, const asc_hip_matrix_desc& c,
               bool ordered, asc_hip::SpmmArgs* m) {
  *m = asc_hip::SpmmArgs{};
  m->a = plan.args.a;
  m->b = b.data;
  m->ldb = b.leading_dimension;
  m->b_row_major = b.layout == ASC_HIP_ROW_MAJOR;
  m->c = const_cast<void*>(c.data);
  m->ldc = c.leading_dimension;
  m->c_row_major = c.layout == ASC_HIP_ROW_MAJOR;
  m->k = b.cols;
  m->alpha = 0.0;
  m->ordered = ordered;
}

asc_hip_status spmm_form(const asc_hip_spmv_options* opt, bool* ordered, bool* panel) {
  const uint8_t alg = opt ? opt->algorithm : ASC_HIP_ALG_AUTO;
  if (alg != ASC_HIP_ALG_AUTO && alg != ASC_HIP_ALG_ORDERED && alg != ASC_HIP_ALG_PANEL)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "SpMM takes ASC_HIP_ALG_AUTO (plain), ASC_HIP_ALG_PANEL or ASC_HIP_ALG_ORDERED");
  if (opt && opt->accumulate != ASC_HIP_ACC_PLAIN)
    return fail(ASC_HIP_UNSUPPORTED, "SpMM accumulates in the storage precision only");
  *ordered = alg == ASC_HIP_ALG_ORDERED;
  *panel = alg == ASC_HIP_ALG_PANEL;
  return ok();
}

// The stream-staged column-panel analysis the context holds for `a` (the
// plain SpMM's ASC_HIP_ALG_PANEL form runs on its segment tables).
bool find_stream_plan(asc_hip_context* ctx, const asc_hip_sparse_desc& a,
                      asc_hip::PanelPlan* out) {
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  for (const auto& p : ctx->plans) {
    if (p.offsets == a.outer_offsets && p.indices == a.inner_indices && p.rows == a.rows &&
        p.cols == a.cols && p.nnz == a.nnz && p.element_kind == a.element_kind &&
        p.algorithm == ASC_HIP_ALG_PANEL && p.panel.stream && p.panel.seg_src != nullptr) {
      *out = p.panel;
      return true;
    }
  }
  return false;
}

// Whether ASC_HIP_ALG_AUTO takes the SpMM's panel form on a stream-staged
// analysis: only when its listed rows are long against the panel count, so
// that their (row, panel) segments hold several entries each.  On the
// 640M-entry Zipf matrix (256 panels, k = 32) the panel form took 17.5 /
// 18.3 / 19.9 ms with rows of at least 16 / 8 / 4 entries per panel listed,
// against 24.5 ms for the plain form; an analysis listing every row (the
// stream form's default) gives segments of one or two entries.
constexpr int64_t kSpmmAutoPanelEntries = 8;

bool spmm_panel_pays(const asc_hip::PanelPlan& pp) {
  return pp.nlong > 0 && pp.long_min >= kSpmmAutoPanelEntries * pp.panels;
}

asc_hip_status prepare_spmm(asc_hip_context* ctx, uint8_t transpose,
                            const asc_hip_sparse_desc* a, const asc_hip_matrix_desc* b,
                            const asc_hip_matrix_desc* c, const asc_hip_spmv_options* opt,
                            SpmvPlan* plan, asc_hip::SpmmArgs* m) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (transpose > ASC_HIP_OP_H) return fail(ASC_HIP_INVALID_ARGUMENT, "unknown transpose");
  bool ordered = false, panel = false;
  if ((st = spmm_form(opt, &ordered, &panel)).code) return st;
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "SpMM requires CSR or CSC");
  if (a->element_kind != ASC_HIP_F32 && a->element_kind != ASC_HIP_F64)
    return fail(ASC_HIP_UNSUPPORTED, "SpMM supports real FP32/FP64 storage");
  const int64_t inner = transpose == ASC_HIP_OP_N ? a->cols : a->rows;
  const int64_t rows = transpose == ASC_HIP_OP_N ? a->rows : a->cols;
  if (b == nullptr || c == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "dense descriptor is null");
  if ((st = check_matrix(ctx, b, a->element_kind, inner, -1, false, "input")).code) return st;
  if ((st = check_matrix(ctx, c, a->element_kind, rows, b->cols, true, "output")).code) return st;
  const Span cs = matrix_span(*c);
  if (overlaps(matrix_span(*b), cs) || sparse_overlaps(*a, cs))
    return fail(ASC_HIP_INVALID_ARGUMENT, "Sparse matrix-matrix input and output overlap");
  if ((st = plan_row_view(*a, transpose, plan)).code) return st;
  spmm_args(*plan, *b, *c, ordered, m);
  if (panel) {
    if (transpose != ASC_HIP_OP_N || a->format != ASC_HIP_CSR)
      return fail(ASC_HIP_UNSUPPORTED, "ASC_HIP_ALG_PANEL SpMM takes op(A) = A of a CSR matrix");
    if (!asc_hip::spmm_plain_applies(*m))
      return fail(ASC_HIP_UNSUPPORTED,
                  "ASC_HIP_ALG_PANEL SpMM needs the plain form's operands: row-major B and C, "
                  "k <= 32, k and both leading dimensions multiples of 4, B and C aligned to "
                  "4 elements");
    if (!find_stream_plan(ctx, *a, &plan->spmm_panel))
      return fail(ASC_HIP_INVALID_ARGUMENT,
                  "ASC_HIP_ALG_PANEL SpMM needs a stream-staged column-panel analysis of this "
                  "matrix (asc_sparse_hip_csr_spmv_analyze with ASC_HIP_STAGING_STREAM)");
    if (panel_overlaps(plan->spmm_panel, cs))
      return fail(ASC_HIP_INVALID_ARGUMENT, "SpMM output overlaps the column-panel analysis");
    m->panel = &plan->spmm_panel;
  } else if (transpose == ASC_HIP_OP_N && a->format == ASC_HIP_CSR &&
             asc_hip::spmm_plain_applies(*m) && find_stream_plan(ctx, *a, &plan->spmm_panel) &&
             spmm_panel_pays(plan->spmm_panel) && !panel_overlaps(plan->spmm_panel, cs)) {
    // AUTO or ORDERED on a matrix whose stream-staged analysis lists long
    // rows only.  The ordered form walks the panels in order and needs no
    // workspace (the same bits as without the analysis); AUTO's panel form
    // keeps partial sums in the workspace and is dropped when it is short.
    m->panel = &plan->spmm_panel;
    plan->spmm_panel_optional = !ordered;
  }
  spmm_plain_plan(m, elem_size(a->element_kind));
  plan->spmv_ws = align256(asc_hip::spmm_workspace_bytes(*m));
  plan->total = transpose_prefix(*plan) + plan->spmv_ws;
  plan->noop = rows == 0 || b->cols == 0;
  return ok();
}

}  // namespace

asc_hip_status asc_sparse_hip_standard_spmm_workspace_size(
    asc_hip_context* ctx, uint8_t transpose, const asc_hip_sparse_desc* a,
    const asc_hip_matrix_desc* b, const asc_hip_matrix_desc* c,
    const asc_hip_spmv_options* opt, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  SpmvPlan plan;
  asc_hip::SpmmArgs m;
  asc_hip_status st = prepare_spmm(ctx, transpose, a, b, c, opt, &plan, &m);
  if (st.code) return st;
  *out = plan.total;
  return ok();
}

asc_hip_status asc_sparse_hip_standard_spmm(asc_hip_context* ctx, uint8_t transpose,
                                            asc_hip_scalar alpha,
                                            const asc_hip_sparse_desc* a,
                                            const asc_hip_matrix_desc* b,
                                            const asc_hip_matrix_desc* c, void* ws,
                                            size_t ws_bytes, const asc_hip_spmv_options* opt) {
  SpmvPlan plan;
  asc_hip::SpmmArgs m;
  asc_hip_status st = prepare_spmm(ctx, transpose, a, b, c, opt, &plan, &m);
  if (st.code) return st;
  // The passes' split table is an optimization, and so is the panel form
  // when AUTO took it: a workspace too small for them runs the plain kernel
  // over every row, then in one pass.  An explicit ASC_HIP_ALG_PANEL keeps
  // its partial sums as the required part.
  const auto needed = [&] {
    return transpose_prefix(plan) + align256(asc_hip::spmm_workspace_bytes(m));
  };
  if ((ws_bytes < plan.total || ws == nullptr) && plan.spmm_panel_optional) {
    m.panel = nullptr;
    plan.total = needed();
  }
  if (ws_bytes < plan.total || ws == nullptr) {
    m.panel_cols = 0;
    plan.total = needed();
  }
  if ((st = check_workspace(ctx, ws, ws_bytes, plan.total,
                            operand_spans(*a, {matrix_span(*b), matrix_span(*c)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  if (alpha.real == 0.0 || plan.noop) return ok();  // blas.h:1730-1732
  DeviceGuard guard(ctx->device);
  if (plan.transposed) {
    const hipError_t e = run_transpose(&plan, static_cast<char*>(ws), ctx->stream);
    if (e != hipSuccess) return finish_launch(ctx, e, "transpose");
  }
  m.a = plan.args.a;  // the transposed operand when op(A) needed one
  m.alpha = alpha.real;
  return finish_launch(ctx,
                       asc_hip::spmm_launch(m, static_cast<char*>(ws) + transpose_prefix(plan),
                                            ctx->stream),
                       "SpMM launch");
}

// ---- conversions ---------------------------------------------------------------

namespace {

asc_hip_status prepare_coo_convert(asc_hip_context* ctx, const asc_hip_sparse_desc* src,
                                   uint8_t out_format, uint8_t index_bytes, size_t* ws) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, src, "source")).code) return st;
  if (src->format != ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "source must be canonical COO");
  if (out_format != ASC_HIP_CSR && out_format != ASC_HIP_CSC)
    return fail(ASC_HIP_INVALID_ARGUMENT, "target format must be CSR or CSC");
  if (index_bytes != 4 && index_bytes != 8) return fail(ASC_HIP_INVALID_ARGUMENT, "index width must be 4 or 8 bytes");
  const int64_t inner = out_format == ASC_HIP_CSR ? src->cols : src->rows;
  if (index_bytes == 4 && inner > std::numeric_limits<int32_t>::max())
    return fail(ASC_HIP_OVERFLOW, "int32 indices cannot address the inner extent");
  const bool csc = out_format == ASC_HIP_CSC;
  if (csc && (src->nnz > kU32Max || src->rows > kU32Max || src->cols > kU32Max))
    return fail(ASC_HIP_UNSUPPORTED, "device COO->CSC supports extents and nnz below 2^32");
  *ws = asc_hip::coo_convert_workspace_bytes(src->rows, src->cols, src->nnz, csc,
                                             static_cast<Elem>(src->element_kind));
  return ok();
}

asc_hip_status check_outputs(const asc_hip_context* ctx, const asc_hip_sparse_desc& src,
                             int64_t out_outer, uint8_t index_bytes, void* off,
                             void* idx, void* val) {
  asc_hip_status st;
  const size_t ob = sizeof(int64_t) * (out_outer + 1);
  const size_t ib = static_cast<size_t>(index_bytes) * src.nnz;
  const size_t vb = elem_size(src.element_kind) * src.nnz;
  if ((st = check_device_pointer(ctx, off, "offsets destination", ob)).code) return st;
  if (src.nnz > 0) {
    if ((st = check_device_pointer(ctx, idx, "indices destination", ib)).code) return st;
    if ((st = check_device_pointer(ctx, val, "values destination", vb)).code) return st;
  }
  const Span so = span_of(off, ob), si = span_of(idx, ib), sv = span_of(val, vb);
  if (overlaps(so, si) || overlaps(so, sv) || overlaps(si, sv))
    return fail(ASC_HIP_INVALID_ARGUMENT, "conversion destinations overlap");
  if (sparse_overlaps(src, so) || sparse_overlaps(src, si) || sparse_overlaps(src, sv))
    return fail(ASC_HIP_INVALID_ARGUMENT, "conversion destination overlaps the source");
  return ok();
}

}  // namespace

asc_hip_status asc_sparse_hip_convert_coordinate_workspace_size(
    asc_hip_context* ctx, const asc_hip_sparse_desc* src, uint8_t out_format,
    uint8_t index_bytes, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  return prepare_coo_convert(ctx, src, out_format, index_bytes, out);
}

asc_hip_status asc_sparse_hip_convert_coordinate(
    asc_hip_context* ctx, const asc_hip_sparse_desc* src, uint8_t out_format,
    uint8_t index_bytes, void* off, void* idx, void* val, asc_hip_sparse_desc* out,
    void* ws, size_t ws_bytes) {
  size_t need = 0;
  asc_hip_status st = prepare_coo_convert(ctx, src, out_format, index_bytes, &need);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  const bool csc = out_format == ASC_HIP_CSC;
  const int64_t out_outer = csc ? src->cols : src->rows;
  if ((st = check_outputs(ctx, *src, out_outer, index_bytes, off, idx, val)).code) return st;
  const size_t ob = sizeof(int64_t) * (out_outer + 1);
  if ((st = check_workspace(ctx, ws, ws_bytes, need,
                            operand_spans(*src, {span_of(off, ob),
                                                 span_of(idx, index_bytes * src->nnz),
                                                 span_of(val, elem_size(src->element_kind) * src->nnz)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  const hipError_t e = asc_hip::coo_convert_launch(
      static_cast<const int64_t*>(src->outer_offsets), src->values,
      static_cast<Elem>(src->element_kind), src->rows, src->cols, src->nnz, csc,
      static_cast<int64_t*>(off), idx, index_bytes, val, ws, ctx->stream);
  if ((st = finish_launch(ctx, e, "COO conversion")).code) return st;
  *out = *src;
  out->format = out_format;
  out->outer_offsets = off;
  out->inner_indices = idx;
  out->values = val;
  out->index_bytes = index_bytes;
  out->offset_bytes = 8;
  out->trusted = 1;
  return ok();
}

// ---- COO finalize ------------------------------------------------------------

namespace {

asc_hip_status check_triplets(asc_hip_context* ctx, const asc_hip_sparse_desc* t) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (t == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "triplet descriptor is null");
  if (t->format != ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "finalize takes COO triplets");
  if (t->element_kind != ASC_HIP_F32 && t->element_kind != ASC_HIP_F64)
    return fail(ASC_HIP_UNSUPPORTED, "finalize supports FP32/FP64 values");
  if (t->rows < 0 || t->cols < 0 || t->nnz < 0)
    return fail(ASC_HIP_SHAPE, "triplets: negative extent or count");
  if (t->rows > kU32Max || t->cols > kU32Max || t->nnz > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "device finalize supports extents and counts below 2^32");
  if (t->memory_space != ASC_HIP_DEVICE && t->memory_space != ASC_HIP_MANAGED)
    return fail(ASC_HIP_MEMORY_ACCESS, "triplets: operand is not device memory");
  if (t->nnz > 0) {
    if (t->outer_offsets == nullptr || t->values == nullptr)
      return fail(ASC_HIP_INVALID_ARGUMENT, "triplets: arrays are null");
    const size_t n = static_cast<size_t>(t->nnz);
    if ((st = check_device_pointer(ctx, t->outer_offsets, "triplets", 16 * n)).code) return st;
    if ((st = check_device_pointer(ctx, t->values, "triplets", elem_size(t->element_kind) * n)).code)
      return st;
  }
  return ok();
}

}  // namespace

asc_hip_status asc_sparse_hip_coo_finalize_workspace_size(asc_hip_context* ctx,
                                                          const asc_hip_sparse_desc* t,
                                                          size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  asc_hip_status st = check_triplets(ctx, t);
  if (st.code) return st;
  *out = asc_hip::coo_finalize_workspace_bytes(t->rows, t->cols, t->nnz,
                                               static_cast<Elem>(t->element_kind));
  return ok();
}

asc_hip_status asc_sparse_hip_coo_finalize(asc_hip_context* ctx, const asc_hip_sparse_desc* t,
                                           uint8_t duplicate_policy, uint8_t zero_policy,
                                           void* out_coords, void* out_vals,
                                           asc_hip_sparse_desc* out, void* ws, size_t ws_bytes) {
  size_t need = 0;
  asc_hip_status st = asc_sparse_hip_coo_finalize_workspace_size(ctx, t, &need);
  if (st.code) return st;
  if (duplicate_policy > ASC_HIP_DUPLICATE_SUM)
    return fail(ASC_HIP_INVALID_ARGUMENT, "Unknown sparse duplicate policy");
  if (zero_policy > ASC_HIP_ZERO_DROP)
    return fail(ASC_HIP_INVALID_ARGUMENT, "Unknown sparse explicit-zero policy");
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  const size_t n = static_cast<size_t>(t->nnz);
  const size_t esz = elem_size(t->element_kind);
  if (n > 0) {
    if (out_coords == nullptr || out_vals == nullptr)
      return fail(ASC_HIP_INVALID_ARGUMENT, "output arrays are null");
    if ((st = check_device_pointer(ctx, out_coords, "output", 16 * n)).code) return st;
    if ((st = check_device_pointer(ctx, out_vals, "output", esz * n)).code) return st;
  }
  const Span oc = span_of(out_coords, 16 * n), ov = span_of(out_vals, esz * n);
  const Span ic = span_of(t->outer_offsets, 16 * n), iv = span_of(t->values, esz * n);
  if (overlaps(oc, ov) || overlaps(oc, ic) || overlaps(oc, iv) || overlaps(ov, ic) || overlaps(ov, iv))
    return fail(ASC_HIP_INVALID_ARGUMENT, "finalize outputs overlap each other or the triplets");
  if ((st = check_workspace(ctx, ws, ws_bytes, need, {oc, ov, ic, iv})).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  if (guard.error != hipSuccess) return fail_hip(guard.error, "hipSetDevice");
  int64_t count = 0;
  int failure = 0;
  const hipError_t e = asc_hip::coo_finalize_launch(
      static_cast<const int64_t*>(t->outer_offsets), t->values,
      static_cast<Elem>(t->element_kind), t->rows, t->cols, t->nnz,
      duplicate_policy == ASC_HIP_DUPLICATE_SUM, zero_policy == ASC_HIP_ZERO_DROP,
      static_cast<int64_t*>(out_coords), out_vals, &count, &failure, ws, ctx->stream);
  if ((st = finish_launch(ctx, e, "COO finalize")).code) return st;
  if (failure == 1) return fail(ASC_HIP_INDEX, "Sparse coordinate is outside the shape");
  if (failure == 2) return fail(ASC_HIP_INVALID_ARGUMENT, "Duplicate sparse coordinates were rejected");
  *out = *t;
  out->outer_offsets = out_coords;
  out->inner_indices = nullptr;
  out->values = out_vals;
  out->nnz = count;
  out->index_bytes = 8;
  out->offset_bytes = 8;
  out->memory_space = ASC_HIP_DEVICE;
  out->trusted = 1;
  return ok();
}

asc_hip_status asc_sparse_hip_convert_compressed_workspace_size(
    asc_hip_context* ctx, const asc_hip_sparse_desc* src, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, src, "source")).code) return st;
  if (src->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "source must be CSR or CSC");
  if (src->nnz > kU32Max || src->rows > kU32Max || src->cols > kU32Max)
    return fail(ASC_HIP_UNSUPPORTED, "device transpose supports extents and nnz below 2^32");
  if (src->index_bytes == 4 && outer_of(*src) > std::numeric_limits<int32_t>::max())
    return fail(ASC_HIP_OVERFLOW, "int32 indices cannot address the transposed inner extent");
  *out = asc_hip::transpose_workspace_bytes(as_compressed(*src, false));
  return ok();
}

asc_hip_status asc_sparse_hip_convert_compressed(asc_hip_context* ctx,
                                                 const asc_hip_sparse_desc* src,
                                                 void* off, void* idx, void* val,
                                                 asc_hip_sparse_desc* out, void* ws,
                                                 size_t ws_bytes) {
  size_t need = 0;
  asc_hip_status st = asc_sparse_hip_convert_compressed_workspace_size(ctx, src, &need);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  const int64_t out_outer = inner_of(*src);
  if ((st = check_outputs(ctx, *src, out_outer, src->index_bytes, off, idx, val)).code) return st;
  const size_t ob = sizeof(int64_t) * (out_outer + 1);
  if ((st = check_workspace(ctx, ws, ws_bytes, need,
                            operand_spans(*src, {span_of(off, ob),
                                                 span_of(idx, src->index_bytes * src->nnz),
                                                 span_of(val, elem_size(src->element_kind) * src->nnz)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  const hipError_t e = asc_hip::transpose_launch(as_compressed(*src, false),
                                                 static_cast<int64_t*>(off), idx,
                                                 src->index_bytes, val, ws, ctx->stream);
  if ((st = finish_launch(ctx, e, "compressed conversion")).code) return st;
  *out = *src;
  out->format = src->format == ASC_HIP_CSR ? ASC_HIP_CSC : ASC_HIP_CSR;
  out->outer_offsets = off;
  out->inner_indices = idx;
  out->values = val;
  out->offset_bytes = 8;
  out->trusted = 1;
  return ok();
}

// ---- compressed -> canonical COO ------------------------------------------------

namespace {

// CSC sources go through a CSR transpose in the workspace first.
struct ToCooPlan {
  bool via_transpose = false;
  size_t t_off = 0, t_idx = 0, t_val = 0, t_ws = 0, total = 0;
};

asc_hip_status prepare_to_coo(asc_hip_context* ctx, const asc_hip_sparse_desc* src,
                              ToCooPlan* plan) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, src, "source")).code) return st;
  if (src->format == ASC_HIP_COO) return fail(ASC_HIP_INVALID_ARGUMENT, "source must be CSR or CSC");
  if (src->format == ASC_HIP_CSC) {
    if (src->nnz > kU32Max || src->rows > kU32Max || src->cols > kU32Max)
      return fail(ASC_HIP_UNSUPPORTED, "device transpose supports extents and nnz below 2^32");
    plan->via_transpose = true;
    plan->t_off = align256(sizeof(int64_t) * (src->rows + 1));
    plan->t_idx = align256(static_cast<size_t>(src->index_bytes) * src->nnz);
    plan->t_val = align256(elem_size(src->element_kind) * src->nnz);
    plan->t_ws = align256(asc_hip::transpose_workspace_bytes(as_compressed(*src, false)));
    plan->total = plan->t_off + plan->t_idx + plan->t_val + plan->t_ws;
  }
  return ok();
}

}  // namespace

asc_hip_status asc_sparse_hip_convert_to_coordinate_workspace_size(
    asc_hip_context* ctx, const asc_hip_sparse_desc* src, size_t* out) {
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  ToCooPlan plan;
  asc_hip_status st = prepare_to_coo(ctx, src, &plan);
  if (st.code) return st;
  *out = plan.total;
  return ok();
}

asc_hip_status asc_sparse_hip_convert_to_coordinate(asc_hip_context* ctx,
                                                    const asc_hip_sparse_desc* src,
                                                    void* d_coords, void* d_values,
                                                    asc_hip_sparse_desc* out, void* ws,
                                                    size_t ws_bytes) {
  ToCooPlan plan;
  asc_hip_status st = prepare_to_coo(ctx, src, &plan);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output descriptor is null");
  const size_t cb = 16 * static_cast<size_t>(src->nnz);
  const size_t vb = elem_size(src->element_kind) * src->nnz;
  if (src->nnz > 0) {
    if ((st = check_device_pointer(ctx, d_coords, "coordinates destination", cb)).code) return st;
    if ((st = check_device_pointer(ctx, d_values, "values destination", vb)).code) return st;
    if (overlaps(span_of(d_coords, cb), span_of(d_values, vb)) ||
        sparse_overlaps(*src, span_of(d_coords, cb)) || sparse_overlaps(*src, span_of(d_values, vb)))
      return fail(ASC_HIP_INVALID_ARGUMENT, "conversion destinations overlap");
  }
  if ((st = check_workspace(ctx, ws, ws_bytes, plan.total,
                            operand_spans(*src, {span_of(d_coords, cb), span_of(d_values, vb)}))).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  asc_hip::Compressed csr = as_compressed(*src, false);
  hipError_t e = hipSuccess;
  if (plan.via_transpose) {
    char* w = static_cast<char*>(ws);
    int64_t* off = reinterpret_cast<int64_t*>(w);
    void* idx = w + plan.t_off;
    void* val = w + plan.t_off + plan.t_idx;
    e = asc_hip::transpose_launch(csr, off, idx, src->index_bytes, val,
                                  w + plan.t_off + plan.t_idx + plan.t_val, ctx->stream);
    csr.offsets = off;
    csr.indices = idx;
    csr.values = val;
    csr.outer = src->rows;
    csr.inner = src->cols;
  }
  if (e == hipSuccess)
    e = asc_hip::csr_to_coo_launch(csr, static_cast<int64_t*>(d_coords), d_values, ctx->stream);
  if ((st = finish_launch(ctx, e, "conversion to COO")).code) return st;
  *out = *src;
  out->format = ASC_HIP_COO;
  out->outer_offsets = d_coords;
  out->inner_indices = nullptr;
  out->values = d_values;
  out->index_bytes = 8;
  out->offset_bytes = 8;
  out->trusted = 1;
  return ok();
}

}  // extern "C"
