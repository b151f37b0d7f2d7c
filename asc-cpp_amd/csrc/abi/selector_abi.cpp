// C ABI of the kernel-variant machinery: the exhaustive SpMV autotuner, the
// per-context tuned-plan cache, GPU sparsity features and the learned
// selector's MLP (forward, gradient, Adam) and inference.
//
// The reference has no counterpart (its CUDA facet always calls one
// cuSPARSE algorithm, src/sparse/cuda/operations.cc:1004-1040); the plan
// cache realizes the per-context tuned-plan state the facet's threading
// contract allows (one context per thread, docs/modules/sparse.md:440-442).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using namespace asc_hip_abi;

namespace {

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

asc_hip_status check_csr_for_features(asc_hip_context* ctx, const asc_hip_sparse_desc* a) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if ((st = check_device_sparse(ctx, a, "matrix")).code) return st;
  if (a->format != ASC_HIP_CSR) return fail(ASC_HIP_INVALID_ARGUMENT, "features need a CSR matrix");
  return ok();
}

SpanList operand_spans_for_select(const asc_hip_sparse_desc& a) {
  const SparseSpans s = sparse_spans(a);
  return {s.first, s.second, s.values};
}

size_t select_ws(int32_t k) {
  return align256(asc_hip::features_workspace_bytes()) + align256(sizeof(float) * 32) +
         align256(sizeof(float) * asc_hip::mlp_scratch_floats(1, k)) + align256(sizeof(float) * k);
}

}  // namespace

extern "C" {

asc_hip_status asc_sparse_hip_plan_set(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                       uint8_t algorithm) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (a == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "matrix descriptor is null");
  if (algorithm < ASC_HIP_ALG_ORDERED || algorithm >= ASC_HIP_ALG_COUNT)
    return fail(ASC_HIP_INVALID_ARGUMENT, "unknown SpMV algorithm");
  if (algorithm == ASC_HIP_ALG_PANEL)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "ASC_HIP_ALG_PANEL plans are made by asc_sparse_hip_csr_spmv_analyze");
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  for (auto& p : ctx->plans) {
    if (p.offsets == a->outer_offsets && p.indices == a->inner_indices && p.rows == a->rows &&
        p.cols == a->cols && p.nnz == a->nnz && p.element_kind == a->element_kind) {
      p.algorithm = algorithm;
      p.panel = asc_hip::PanelPlan{};
      return ok();
    }
  }
  ctx->plans.push_back({a->outer_offsets, a->inner_indices, a->rows, a->cols, a->nnz,
                        a->element_kind, algorithm, asc_hip::PanelPlan{}});
  return ok();
}

asc_hip_status asc_sparse_hip_plan_get(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                       uint8_t* out_algorithm) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (a == nullptr || out_algorithm == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  *out_algorithm = ASC_HIP_ALG_AUTO;
  for (const auto& p : ctx->plans)
    if (p.offsets == a->outer_offsets && p.indices == a->inner_indices && p.rows == a->rows &&
        p.cols == a->cols && p.nnz == a->nnz && p.element_kind == a->element_kind)
      *out_algorithm = p.algorithm;
  return ok();
}

asc_hip_status asc_sparse_hip_plan_clear(asc_hip_context* ctx) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  std::lock_guard<std::mutex> lock(ctx->plan_mutex);
  ctx->plans.clear();
  ctx->trsv_schedules.clear();
  return ok();
}

asc_hip_status asc_sparse_hip_csr_spmv_autotune(asc_hip_context* ctx,
                                                const asc_hip_sparse_desc* a,
                                                const asc_hip_vector_desc* x,
                                                const asc_hip_vector_desc* y, void* ws,
                                                size_t ws_bytes, int32_t reps,
                                                float* times_ms, uint8_t* out_alg) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (reps < 1) return fail(ASC_HIP_INVALID_ARGUMENT, "reps must be positive");
  if (out_alg == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  // Validate once with the variant that needs the most workspace.
  asc_hip_spmv_options opt{ASC_HIP_ALG_MERGE, ASC_HIP_ACC_PLAIN, {}};
  size_t need = 0;
  if ((st = asc_sparse_hip_csr_spmv_workspace_size(ctx, a, x, y, &opt, &need)).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err != hipSuccess) return fail_hip(err, "hipEventCreate");
  float best = -1.0f;
  uint8_t best_alg = ASC_HIP_ALG_MERGE;
  for (uint8_t alg = ASC_HIP_ALG_MERGE; alg <= ASC_HIP_ALG_VECTOR64 && st.code == 0; ++alg) {
    opt.algorithm = alg;
    st = asc_sparse_hip_csr_spmv(ctx, 1.0, a, x, 0.0, y, ws, ws_bytes, &opt);  // warm-up
    if (st.code) break;
    err = hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps && err == hipSuccess && st.code == 0; ++r)
      st = asc_sparse_hip_csr_spmv(ctx, 1.0, a, x, 0.0, y, ws, ws_bytes, &opt);
    if (st.code) break;
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err != hipSuccess) {
      st = finish_launch(ctx, err, "autotune timing");
      break;
    }
    ms /= static_cast<float>(reps);
    if (times_ms) times_ms[alg] = ms;
    if (best < 0.0f || ms < best) {
      best = ms;
      best_alg = alg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st.code) return st;
  *out_alg = best_alg;
  return asc_sparse_hip_plan_set(ctx, a, best_alg);
}

asc_hip_status asc_sparse_hip_csr_features_workspace_size(asc_hip_context* ctx, size_t* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  *out = asc_hip::features_workspace_bytes();
  return ok();
}

asc_hip_status asc_sparse_hip_csr_features(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                           float* device_out16, void* ws, size_t ws_bytes) {
  asc_hip_status st = check_csr_for_features(ctx, a);
  if (st.code) return st;
  if ((st = check_device_pointer(ctx, device_out16, "features output")).code) return st;
  if ((st = check_workspace(ctx, ws, ws_bytes, asc_hip::features_workspace_bytes(),
                            {span_of(device_out16, 64)})).code)
    return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::features_launch(as_compressed(*a, false), device_out16, ws,
                                                     ctx->stream),
                       "features");
}

asc_hip_status asc_sparse_hip_mlp_sizes(int32_t k, int64_t batch, size_t* param_floats,
                                        size_t* scratch_bytes) {
  if (k < 1 || k > 64 || batch < 1) return fail(ASC_HIP_INVALID_ARGUMENT, "bad MLP shape");
  if (param_floats) *param_floats = asc_hip::mlp_param_count(k);
  if (scratch_bytes) *scratch_bytes = sizeof(float) * asc_hip::mlp_scratch_floats(batch, k);
  return ok();
}

asc_hip_status asc_sparse_hip_mlp_forward(asc_hip_context* ctx, int32_t k, const float* params,
                                          int64_t batch, const float* x, float* y,
                                          void* scratch, size_t scratch_bytes) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  size_t need = 0;
  if ((st = asc_sparse_hip_mlp_sizes(k, batch, nullptr, &need)).code) return st;
  if (scratch == nullptr || scratch_bytes < need)
    return fail(ASC_HIP_INVALID_ARGUMENT, "MLP scratch is smaller than required");
  for (const void* p : {static_cast<const void*>(params), static_cast<const void*>(x),
                        static_cast<const void*>(y), static_cast<const void*>(scratch)})
    if ((st = check_device_pointer(ctx, p, "MLP operand")).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::mlp_forward_launch(params, k, batch, x, y,
                                                        static_cast<float*>(scratch), ctx->stream),
                       "MLP forward");
}

asc_hip_status asc_sparse_hip_mlp_gradient(asc_hip_context* ctx, int32_t k, const float* params,
                                           int64_t batch, const float* x, const float* target,
                                           float* grads, float* loss, void* scratch,
                                           size_t scratch_bytes) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  size_t need = 0;
  if ((st = asc_sparse_hip_mlp_sizes(k, batch, nullptr, &need)).code) return st;
  if (scratch == nullptr || scratch_bytes < need)
    return fail(ASC_HIP_INVALID_ARGUMENT, "MLP scratch is smaller than required");
  for (const void* p : {static_cast<const void*>(params), static_cast<const void*>(x),
                        static_cast<const void*>(target), static_cast<const void*>(grads),
                        static_cast<const void*>(loss), static_cast<const void*>(scratch)})
    if ((st = check_device_pointer(ctx, p, "MLP operand")).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::mlp_grad_launch(params, k, batch, x, target, grads,
                                                     static_cast<float*>(scratch), loss,
                                                     ctx->stream),
                       "MLP gradient");
}

asc_hip_status asc_sparse_hip_mlp_adam(asc_hip_context* ctx, float* params, const float* grads,
                                       float* m, float* v, int64_t n, float lr, float beta1,
                                       float beta2, float eps, int64_t step) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (n < 1 || step < 1 || !(lr > 0.0f)) return fail(ASC_HIP_INVALID_ARGUMENT, "bad Adam arguments");
  for (const void* p : {static_cast<const void*>(params), static_cast<const void*>(grads),
                        static_cast<const void*>(m), static_cast<const void*>(v)})
    if ((st = check_device_pointer(ctx, p, "Adam operand")).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::mlp_adam_launch(params, grads, m, v, n, lr, beta1, beta2,
                                                     eps, step, ctx->stream),
                       "Adam");
}

asc_hip_status asc_sparse_hip_normalize_features(asc_hip_context* ctx, const float* x,
                                                 const float* mean, const float* stddev,
                                                 int64_t n, float* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (n < 1) return fail(ASC_HIP_INVALID_ARGUMENT, "bad row count");
  for (const void* p : {static_cast<const void*>(x), static_cast<const void*>(mean),
                        static_cast<const void*>(stddev), static_cast<const void*>(out)})
    if ((st = check_device_pointer(ctx, p, "normalization operand")).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  DeviceGuard guard(ctx->device);
  return finish_launch(ctx, asc_hip::normalize_launch(x, mean, stddev, n, out, ctx->stream),
                       "normalize");
}

asc_hip_status asc_sparse_hip_select_workspace_size(asc_hip_context* ctx,
                                                    const asc_hip_selector* sel, size_t* out) {
  asc_hip_status st = check_context(ctx);
  if (st.code) return st;
  if (sel == nullptr || out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "null argument");
  if (sel->outputs != ASC_HIP_SELECT_DIRECT_OUTPUTS && sel->outputs != ASC_HIP_SELECT_PLAN_OUTPUTS)
    return fail(ASC_HIP_INVALID_ARGUMENT,
                "selector outputs are the 8 direct kernels, or those and the 3 panel plans");
  *out = select_ws(sel->outputs);
  return ok();
}

asc_hip_status asc_sparse_hip_select(asc_hip_context* ctx, const asc_hip_sparse_desc* a,
                                     const asc_hip_selector* sel, void* ws, size_t ws_bytes,
                                     asc_hip_selection* out, float* out_scores) {
  size_t need = 0;
  asc_hip_status st = asc_sparse_hip_select_workspace_size(ctx, sel, &need);
  if (st.code) return st;
  if ((st = check_csr_for_features(ctx, a)).code) return st;
  if (out == nullptr) return fail(ASC_HIP_INVALID_ARGUMENT, "output pointer is null");
  if ((st = check_device_pointer(ctx, sel->device_model, "selector model")).code) return st;
  if ((st = check_workspace(ctx, ws, ws_bytes, need, operand_spans_for_select(*a))).code) return st;
  if ((st = unavailable_if_validation_only(ctx)).code) return st;
  const int k = sel->outputs;
  char* w = static_cast<char*>(ws);
  void* fws = w;
  w += align256(asc_hip::features_workspace_bytes());
  float* feat = reinterpret_cast<float*>(w);
  float* norm = feat + 16;
  w += align256(sizeof(float) * 32);
  float* scratch = reinterpret_cast<float*>(w);
  w += align256(sizeof(float) * asc_hip::mlp_scratch_floats(1, k));
  float* scores = reinterpret_cast<float*>(w);
  DeviceGuard guard(ctx->device);
  const float* model = sel->device_model;
  hipError_t e = asc_hip::features_launch(as_compressed(*a, false), feat, fws, ctx->stream);
  if (e == hipSuccess) e = asc_hip::normalize_launch(feat, model, model + 16, 1, norm, ctx->stream);
  if (e == hipSuccess)
    e = asc_hip::mlp_forward_launch(model + 32, k, 1, norm, scores, scratch, ctx->stream);
  float host[32 + 64];
  if (e == hipSuccess)
    e = hipMemcpyAsync(host, norm, sizeof(float) * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host + 32, scores, sizeof(float) * k, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return finish_launch(ctx, e, "selector inference");
  // Predicted log-cost regrets: pick the smallest, unless unsure.
  bool in_range = true;
  for (int f = 0; f < 16; ++f)
    if (!(std::fabs(host[f]) <= sel->clip)) in_range = false;
  int bestk = 0, best_direct = 0;
  for (int j = 1; j < k; ++j) {
    if (host[32 + j] < host[32 + bestk]) bestk = j;
    if (j < ASC_HIP_SELECT_DIRECT_OUTPUTS && host[32 + j] < host[32 + best_direct]) best_direct = j;
  }
  const bool unsure = !in_range || host[32 + 0] - host[32 + bestk] < sel->margin;
  asc_hip_selection pick{};
  pick.fell_back = unsure ? 1 : 0;
  pick.direct_algorithm =
      static_cast<uint8_t>(unsure ? ASC_HIP_ALG_MERGE : ASC_HIP_ALG_MERGE + best_direct);
  if (!unsure && bestk >= ASC_HIP_SELECT_DIRECT_OUTPUTS) {
    // Outputs 8, 9, 10: the L2-, LDS- and stream-staged column-panel plans.
    pick.algorithm = ASC_HIP_ALG_PANEL;
    pick.staging = static_cast<uint8_t>(ASC_HIP_STAGING_L2 + (bestk - ASC_HIP_SELECT_DIRECT_OUTPUTS));
  } else {
    pick.algorithm = pick.direct_algorithm;
    pick.staging = 0;
  }
  if (out_scores)
    for (int j = 0; j < k; ++j) out_scores[j] = host[32 + j];
  *out = pick;
  return asc_sparse_hip_plan_set(ctx, a, pick.direct_algorithm);
}

}  // extern "C"
