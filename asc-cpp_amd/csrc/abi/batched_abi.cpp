// C ABI of the batched SpMV: plan building (validation, work list), the
// persistent launch, hipGraph capture and the per-size-class autotuner.
//
// The reference has no batched entry point; each batch entry follows the
// compat SpMV contract of asc::Spmv(ctx, alpha, CsrView, x, beta, y)
// (include/asc/sparse/blas.h:2031-2081) and is validated like
// asc_sparse_hip_csr_spmv.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using asc_hip::Elem;
using namespace asc_hip_abi;

struct asc_hip_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
};

namespace {

constexpr int64_t kItemNnz = 8192;   // target nonzeros per work item
constexpr int kClasses = 8;          // average-row-length classes [2^c, 2^(c+1))
constexpr int64_t kWorkgroupsPerCu = 8;
constexpr int kChoices = 8;  // tuner candidates: lanes 1..64, then CSR-stream

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int row_class(const asc_hip_sparse_desc& a) {
  const int64_t avg = a.rows > 0 ? a.nnz / a.rows : 0;
  int c = 0;
  while (c < kClasses - 1 && (int64_t(2) << c) <= avg) ++c;
  return c;
}

// Rows per work item: about kItemNnz entries for the lane path, and about
// three quarters of the LDS window for a CSR-stream item (lanes == 0), so
// that rows longer than average rarely overflow it.
int64_t rows_per_item(const asc_hip_sparse_desc& a, int lanes) {
  const int64_t avg = std::max<int64_t>(1, a.rows > 0 ? a.nnz / a.rows : 1);
  const int64_t target = lanes == 0 ? asc_hip::kStreamEntries * 3 / 4 : kItemNnz;
  return std::max<int64_t>(1, target / avg);
}

// Items of the finest cut any class may take (CSR-stream), so plan memory of
// this size serves every choice of lanes.
int64_t count_items(int64_t n, const asc_hip_batch_entry* e) {
  int64_t items = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t rows = e[i].matrix.rows;
    const int64_t per = std::min(rows_per_item(e[i].matrix, 0), rows_per_item(e[i].matrix, 1));
    items += (rows + per - 1) / per;
  }
  return items;
}

asc_hip_status validate_entries(asc_hip_context* ctx, int64_t n, const asc_hip_batch_entry* e,
                                uint8_t* elem, uint8_t* ib) {
  if (n < 0) return fail(ASC_HIP_INVALID_ARGUMENT, "negative batch count");
  if (n > 0 && e == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "batch entries are null");
  if (n > INT32_MAX) return fail(ASC_HIP_OVERFLOW, "batch count exceeds INT32_MAX");
  *elem = n > 0 ? e[0].matrix.element_kind : ASC_HIP_F64;
  *ib = n > 0 ? e[0].matrix.index_bytes : 4;
  std::vector<std::pair<Span, int64_t>> outs, ins;
  outs.reserve(n);
  ins.reserve(4 * n);
  for (int64_t i = 0; i < n; ++i) {
    const asc_hip_batch_entry& b = e[i];
    asc_hip_status st = check_device_sparse(ctx, &b.matrix, "batch matrix");
    if (st.code) return fail(st.code, "entry " + std::to_string(i) + ": " + g_message);
    if (b.matrix.format != ASC_HIP_CSR) return fail(ASC_HIP_INVALID_ARGUMENT, "batch entries must be CSR");
    if (b.matrix.element_kind != ASC_HIP_F64 && b.matrix.element_kind != ASC_HIP_F32)
      return fail(ASC_HIP_UNSUPPORTED, "batched SpMV supports FP32/FP64 storage");
    if (b.matrix.element_kind != *elem || b.matrix.index_bytes != *ib)
      return fail(ASC_HIP_INVALID_ARGUMENT, "batch entries must share element kind and index width");
    if ((st = check_vector(ctx, &b.x, *elem, b.matrix.cols, false, "batch input")).code)
      return fail(st.code, "entry " + std::to_string(i) + ": " + g_message);
    if ((st = check_vector(ctx, &b.y, *elem, b.matrix.rows, true, "batch output")).code)
      return fail(st.code, "entry " + std::to_string(i) + ": " + g_message);
    outs.push_back({vector_span(b.y), i});
    const SparseSpans s = sparse_spans(b.matrix);
    ins.push_back({s.first, i});
    ins.push_back({s.second, i});
    ins.push_back({s.values, i});
    ins.push_back({vector_span(b.x), i});
  }
  // Outputs must be pairwise disjoint and disjoint from every input.
  auto by_begin = [](const auto& a, const auto& b) { return a.first.begin < b.first.begin; };
  std::sort(outs.begin(), outs.end(), by_begin);
  for (size_t i = 1; i < outs.size(); ++i)
    if (overlaps(outs[i - 1].first, outs[i].first))
      return fail(ASC_HIP_INVALID_ARGUMENT, "batch outputs overlap (entries " +
                  std::to_string(outs[i - 1].second) + ", " + std::to_string(outs[i].second) + ")");
  for (const auto& in : ins) {
    if (in.first.begin == in.first.end) continue;
    auto it = std::upper_bound(outs.begin(), outs.end(), in,
                               [](const auto& a, const auto& b) { return a.first.begin < b.first.begin; });
    // candidates: the output starting at or before in.begin, and those after
    if (it != outs.begin() && overlaps(std::prev(it)->first, in.first))
      return fail(ASC_HIP_INVALID_ARGUMENT, "a batch output overlaps an input (entry " +
                  std::to_string(in.second) + ")");
    for (; it != outs.end() && it->first.begin < in.first.end; ++it)
      if (overlaps(it->first, in.first))
        return fail(ASC_HIP_INVALID_ARGUMENT, "a batch output overlaps an input (entry " +
                    std::to_string(in.second) + ")");
  }
  return ok();
}

// Largest persistent grid a plan's range table holds.
constexpr int64_t kMaxGrid = 16384;

size_t plan_bytes(int64_t n, int64_t items) {
  return align256(sizeof(asc_hip::BatchMatrix) * std::max<int64_t>(n, 1)) +
         align256(sizeof(asc_hip::BatchItem) * std::max<int64_t>(items, 1)) +
         align256(sizeof(int64_t) * (kMaxGrid + 1));
}

int64_t persistent_grid(const asc_hip_context* ctx, int64_t items) {
  int cus = 256;
  if (ctx->device >= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && v > 0)
      cus = v;
  }
  return std::max<int64_t>(1, std::min<int64_t>({items, cus * kWorkgroupsPerCu, kMaxGrid}));
}

// Builds host tables for the entries selected by `pick` (all when null) and
// uploads them; fills `plan`.
asc_hip_status build(asc_hip_context* ctx, int64_t n, const asc_hip_batch_entry* e,
                     const uint8_t* lanes, const std::vector<int64_t>* pick, void* mem,
                     size_t bytes, uint8_t elem, uint8_t ib, asc_hip_batch_plan* plan) {
  std::vector<asc_hip::BatchMatrix> mats;
  std::vector<asc_hip::BatchItem> items;
  std::vector<int64_t> weight;  // estimated entries + rows per item
  const int64_t count = pick ? static_cast<int64_t>(pick->size()) : n;
  mats.reserve(count);
  for (int64_t j = 0; j < count; ++j) {
    const asc_hip_batch_entry& b = e[pick ? (*pick)[j] : j];
    asc_hip::BatchMatrix m;
    m.offsets = static_cast<const int64_t*>(b.matrix.outer_offsets);
    m.indices = b.matrix.inner_indices;
    m.values = b.matrix.values;
    m.x = b.x.data;
    m.y = const_cast<void*>(b.y.data);
    m.incx = b.x.stride;
    m.incy = b.y.stride;
    m.alpha = b.alpha;
    m.beta = b.beta;
    mats.push_back(m);
    const int g = lanes[row_class(b.matrix)];
    const int64_t per = rows_per_item(b.matrix, g);
    const int64_t avg = b.matrix.rows > 0 ? b.matrix.nnz / b.matrix.rows : 0;
    for (int64_t r = 0; r < b.matrix.rows; r += per) {
      asc_hip::BatchItem it;
      it.matrix = static_cast<int32_t>(j);
      it.lanes = g;
      it.row_begin = r;
      it.row_end = std::min(b.matrix.rows, r + per);
      items.push_back(it);
      weight.push_back((it.row_end - it.row_begin) * (avg + 1));
    }
  }
  // Contiguous item ranges of about equal weight, one per work-group.
  const int64_t grid = persistent_grid(ctx, static_cast<int64_t>(items.size()));
  std::vector<int64_t> ranges(static_cast<size_t>(grid) + 1, 0);
  {
    int64_t total = 0;
    for (int64_t w : weight) total += w;
    int64_t acc = 0;
    size_t i = 0;
    for (int64_t w = 1; w < grid; ++w) {
      const int64_t goal = (total / grid) * w + (total % grid) * w / grid;
      while (i < items.size() && acc + weight[i] / 2 < goal) acc += weight[i++];
      ranges[static_cast<size_t>(w)] = static_cast<int64_t>(i);
    }
    ranges[static_cast<size_t>(grid)] = static_cast<int64_t>(items.size());
  }
  const size_t need = plan_bytes(count, static_cast<int64_t>(items.size()));
  if (mem == nullptr || bytes < need)
    return fail(ASC_HIP_INVALID_ARGUMENT, "plan memory is smaller than the queried size");
  if (reinterpret_cast<uintptr_t>(mem) % 256 != 0)
    return fail(ASC_HIP_INVALID_ARGUMENT, "plan memory must be 256-byte aligned");
  asc_hip_status st = check_device_pointer(ctx, mem, "plan memory", need);
  if (st.code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  char* base = static_cast<char*>(mem);
  const size_t items_off = align256(sizeof(asc_hip::BatchMatrix) * std::max<int64_t>(count, 1));
  const size_t ranges_off =
      items_off + align256(sizeof(asc_hip::BatchItem) * std::max<size_t>(items.size(), 1));
  hipError_t err = hipSuccess;
  if (!mats.empty())
    err = hipMemcpyAsync(base, mats.data(), sizeof(asc_hip::BatchMatrix) * mats.size(),
                         hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess && !items.empty())
    err = hipMemcpyAsync(base + items_off, items.data(), sizeof(asc_hip::BatchItem) * items.size(),
                         hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(base + ranges_off, ranges.data(), sizeof(int64_t) * ranges.size(),
                         hipMemcpyHostToDevice, ctx->stream);
  const hipError_t sync = hipStreamSynchronize(ctx->stream);
  if (err == hipSuccess) err = sync;
  if (err != hipSuccess) return finish_launch(ctx, err, "plan upload");
  std::memset(plan, 0, sizeof(*plan));
  plan->matrix_table = base;
  plan->item_table = base + items_off;
  plan->matrices = count;
  plan->items = static_cast<int64_t>(items.size());
  plan->grid = grid;
  plan->element_kind = elem;
  plan->index_bytes = ib;
  plan->device = ctx->device;
  return ok();
}

asc_hip::BatchLaunch as_launch(const asc_hip_batch_plan& p) {
  asc_hip::BatchLaunch L;
  L.mats = static_cast<const asc_hip::BatchMatrix*>(p.matrix_table);
  L.items = static_cast<const asc_hip::BatchItem*>(p.item_table);
  L.ranges = reinterpret_cast<const int64_t*>(
      static_cast<const char*>(p.item_table) +
      align256(sizeof(asc_hip::BatchItem) * std::max<int64_t>(p.items, 1)));
  L.nitems = p.items;
  L.grid = p.grid;
  L.elem = static_cast<Elem>(p.element_kind);
  L.index_bytes = p.index_bytes;
  return L;
}

asc_hip_status check_plan(const asc_hip_context* ctx, const asc_hip_batch_plan* p) {
  if (p == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "plan is null");
  if (p->device != ctx->device) return fail(ASC_HIP_INVALID_ARGUMENT, "plan belongs to another device");
  if (p->items > 0 && (p->matrix_table == nullptr || p->item_table == nullptr || p->grid < 1))
    return fail(ASC_HIP_INVALID_STATE, "plan is not built");
  return ok();
}

const uint8_t kDefaultLanes[kClasses] = {1, 2, 4, 8, 16, 32, 64, 64};

}  // namespace

extern "C" {

asc_hip_status asc_sparse_hip_batch_plan_size(asc_hip_context* ctx, int64_t count,
                                              const asc_hip_batch_entry* entries,
                                              size_t* out_bytes) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out_bytes == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  uint8_t elem, ib;
  if ((st = validate_entries(ctx, count, entries, &elem, &ib)).code) return st;
  *out_bytes = plan_bytes(count, count_items(count, entries));
  return ok();
}

asc_hip_status asc_sparse_hip_batch_plan_build(asc_hip_context* ctx, int64_t count,
                                               const asc_hip_batch_entry* entries,
                                               const asc_hip_batch_classes* classes,
                                               void* plan_memory, size_t plan_bytes_given,
                                               asc_hip_batch_plan* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output plan is null");
  uint8_t elem, ib;
  if ((st = validate_entries(ctx, count, entries, &elem, &ib)).code) return st;
  const uint8_t* lanes = classes ? classes->lanes : kDefaultLanes;
  for (int c = 0; c < kClasses; ++c) {
    const uint8_t g = lanes[c];
    if (g > 64 || (g & (g - 1)) != 0)
      return fail(ASC_HIP_INVALID_ARGUMENT,
                  "class lanes must be powers of two in [1, 64], or 0 for CSR-stream");
  }
  return build(ctx, count, entries, lanes, nullptr, plan_memory, plan_bytes_given, elem, ib, out);
}

asc_hip_status asc_sparse_hip_batch_spmv(asc_hip_context* ctx, const asc_hip_batch_plan* plan) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_plan(ctx, plan)).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::batched_spmv_launch(as_launch(*plan), ctx->stream),
                       "batched SpMV launch");
}

asc_hip_status asc_sparse_hip_batch_graph_create(asc_hip_context* ctx,
                                                 const asc_hip_batch_plan* plan,
                                                 asc_hip_graph** out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = nullptr;
  if ((st = check_plan(ctx, plan)).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  auto* g = new (std::nothrow) asc_hip_graph();
  if (g == nullptr) return fail(ASC_HIP_ALLOCATION, "graph allocation failed");
  g->device = ctx->device;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete g;
    return fail_hip(e, "hipStreamBeginCapture");
  }
  const hipError_t launch = asc_hip::batched_spmv_launch(as_launch(*plan), ctx->stream);
  e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (launch != hipSuccess || e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return fail_hip(launch != hipSuccess ? launch : e, "graph capture");
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    return fail_hip(e, "hipGraphInstantiate");
  }
  *out = g;
  return ok();
}

asc_hip_status asc_sparse_hip_graph_launch(asc_hip_context* ctx, asc_hip_graph* g) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (g == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "graph is null");
  if (g->device != ctx->device) return fail(ASC_HIP_INVALID_ARGUMENT, "graph belongs to another device");
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, hipGraphLaunch(g->exec, ctx->stream), "hipGraphLaunch");
}

asc_hip_status asc_sparse_hip_graph_destroy(asc_hip_graph* g) {
  if (g == nullptr) return ok();
  DeviceGuard guard(g->device);
  hipError_t e1 = hipGraphExecDestroy(g->exec);
  hipError_t e2 = hipGraphDestroy(g->graph);
  delete g;
  if (e1 != hipSuccess) return fail_hip(e1, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return fail_hip(e2, "hipGraphDestroy");
  return ok();
}

asc_hip_status asc_sparse_hip_batch_tune(asc_hip_context* ctx, int64_t count,
                                         const asc_hip_batch_entry* entries, void* plan_memory,
                                         size_t plan_bytes_given, int32_t reps,
                                         asc_hip_batch_classes* out, float* class_ms) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output classes are null");
  if (reps < 1) return fail(ASC_HIP_INVALID_ARGUMENT, "reps must be positive");
  uint8_t elem, ib;
  if ((st = validate_entries(ctx, count, entries, &elem, &ib)).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  std::vector<std::vector<int64_t>> members(kClasses);
  for (int64_t i = 0; i < count; ++i) members[row_class(entries[i].matrix)].push_back(i);
  std::memcpy(out->lanes, kDefaultLanes, kClasses);
  DeviceGuard guard(ctx->device);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err != hipSuccess) return fail_hip(err, "hipEventCreate");
  asc_hip_status result = ok();
  for (int c = 0; c < kClasses && result.code == 0; ++c) {
    float best = -1.0f;
    for (int li = 0; li < kChoices; ++li) {
      const uint8_t g = li < 7 ? static_cast<uint8_t>(1 << li) : 0;  // last: CSR-stream
      if (class_ms) class_ms[c * kChoices + li] = -1.0f;
      if (members[c].empty()) continue;
      uint8_t lanes[kClasses];
      std::memcpy(lanes, kDefaultLanes, kClasses);
      lanes[c] = g;
      asc_hip_batch_plan p;
      result = build(ctx, count, entries, lanes, &members[c], plan_memory, plan_bytes_given, elem, ib, &p);
      if (result.code) break;
      const asc_hip::BatchLaunch L = as_launch(p);
      err = asc_hip::batched_spmv_launch(L, ctx->stream);  // warm-up
      if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
      for (int r = 0; r < reps && err == hipSuccess; ++r) err = asc_hip::batched_spmv_launch(L, ctx->stream);
      if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err != hipSuccess) {
        result = finish_launch(ctx, err, "batch tuning");
        break;
      }
      ms /= static_cast<float>(reps);
      if (class_ms) class_ms[c * kChoices + li] = ms;
      if (best < 0.0f || ms < best) {
        best = ms;
        out->lanes[c] = g;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return result;
}

}  // extern "C"
