// Hand-written FP32 kernels of the kernel-selector MLP (16 -> 64 -> 64 -> K):
// forward, backward (mean-squared error on K outputs) and Adam.  Every sum
// runs in a fixed order in one thread, so training is deterministic.
//
// Parameter layout (floats): W1[64][16] b1[64] W2[64][64] b2[64] W3[K][64]
// b3[K].
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void dense_forward_kernel(
    const float* __restrict__ in, int in_dim, const float* __restrict__ w,
    const float* __restrict__ bias, float* __restrict__ out, int out_dim,
    int64_t batch, int relu) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= batch * out_dim) return;
  const int64_t b = t / out_dim;
  const int o = static_cast<int>(t % out_dim);
  float s = bias[o];
  const float* row = in + b * in_dim;
  const float* wr = w + static_cast<int64_t>(o) * in_dim;
  for (int i = 0; i < in_dim; ++i) s = __builtin_fmaf(wr[i], row[i], s);
  out[t] = relu && s < 0.0f ? 0.0f : s;
}

// dY = 2 (Y - T) / (B K); per-block partial of the squared error.
__global__ __launch_bounds__(kBlock) void mse_grad_kernel(
    const float* __restrict__ y, const float* __restrict__ target, int64_t n,
    float* __restrict__ dy, float* __restrict__ partial) {
  __shared__ float s_red[kBlock];
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  float e2 = 0.0f;
  if (t < n) {
    const float d = y[t] - target[t];
    dy[t] = 2.0f * d / static_cast<float>(n);
    e2 = d * d;
  }
  s_red[threadIdx.x] = e2;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s_red[0];
}

__global__ void loss_final_kernel(const float* __restrict__ partial, int nblocks,
                                  int64_t n, float* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; ++b) s += partial[b];
  *loss = s / static_cast<float>(n);
}

// dW[o][i] = sum_b dout[b][o] * in[b][i];  db[o] = sum_b dout[b][o].
__global__ __launch_bounds__(kBlock) void dense_grad_weights_kernel(
    const float* __restrict__ dout, int out_dim, const float* __restrict__ in,
    int in_dim, int64_t batch, float* __restrict__ dw, float* __restrict__ db) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t nw = static_cast<int64_t>(out_dim) * (in_dim + 1);
  if (t >= nw) return;
  const int o = static_cast<int>(t / (in_dim + 1));
  const int i = static_cast<int>(t % (in_dim + 1));
  float s = 0.0f;
  if (i < in_dim) {
    for (int64_t b = 0; b < batch; ++b)
      s = __builtin_fmaf(dout[b * out_dim + o], in[b * in_dim + i], s);
    dw[static_cast<int64_t>(o) * in_dim + i] = s;
  } else {
    for (int64_t b = 0; b < batch; ++b) s += dout[b * out_dim + o];
    db[o] = s;
  }
}

// din[b][i] = [act[b][i] > 0] * sum_o dout[b][o] * w[o][i].
__global__ __launch_bounds__(kBlock) void dense_grad_input_kernel(
    const float* __restrict__ dout, int out_dim, const float* __restrict__ w,
    const float* __restrict__ act, int in_dim, int64_t batch,
    float* __restrict__ din) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= batch * in_dim) return;
  const int64_t b = t / in_dim;
  const int i = static_cast<int>(t % in_dim);
  float s = 0.0f;
  for (int o = 0; o < out_dim; ++o)
    s = __builtin_fmaf(dout[b * out_dim + o], w[static_cast<int64_t>(o) * in_dim + i], s);
  din[t] = act[t] > 0.0f ? s : 0.0f;
}

__global__ __launch_bounds__(kBlock) void adam_kernel(
    float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
    float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
    float bc1, float bc2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.0f - b1) * gi;
  const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

size_t mlp_param_count(int k) { return 64 * 16 + 64 + 64 * 64 + 64 + static_cast<size_t>(k) * 64 + k; }

size_t mlp_scratch_floats(int64_t batch, int k) {
  // H1, H2, Y, dY, dH2, dH1 activations plus loss partials and the loss.
  return static_cast<size_t>(batch) * (64 + 64 + k + k + 64 + 64) +
         static_cast<size_t>((batch * k + kBlock - 1) / kBlock) + 1;
}

struct MlpPtrs {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
};

static MlpPtrs split(const float* p, int k) {
  MlpPtrs r;
  r.w1 = p;
  r.b1 = r.w1 + 64 * 16;
  r.w2 = r.b1 + 64;
  r.b2 = r.w2 + 64 * 64;
  r.w3 = r.b2 + 64;
  r.b3 = r.w3 + k * 64;
  return r;
}

hipError_t mlp_forward_launch(const float* params, int k, int64_t batch, const float* x,
                              float* y, float* scratch, hipStream_t s) {
  const MlpPtrs w = split(params, k);
  float* h1 = scratch;
  float* h2 = h1 + batch * 64;
  dense_forward_kernel<<<blocks(batch * 64), kBlock, 0, s>>>(x, 16, w.w1, w.b1, h1, 64, batch, 1);
  dense_forward_kernel<<<blocks(batch * 64), kBlock, 0, s>>>(h1, 64, w.w2, w.b2, h2, 64, batch, 1);
  dense_forward_kernel<<<blocks(batch * k), kBlock, 0, s>>>(h2, 64, w.w3, w.b3, y, k, batch, 0);
  return hipGetLastError();
}

hipError_t mlp_grad_launch(const float* params, int k, int64_t batch, const float* x,
                           const float* target, float* grads, float* scratch,
                           float* loss, hipStream_t s) {
  const MlpPtrs w = split(params, k);
  float* h1 = scratch;
  float* h2 = h1 + batch * 64;
  float* y = h2 + batch * 64;
  float* dy = y + batch * k;
  float* dh2 = dy + batch * k;
  float* dh1 = dh2 + batch * 64;
  float* partial = dh1 + batch * 64;
  MlpPtrs g = split(grads, k);
  const hipError_t fe = mlp_forward_launch(params, k, batch, x, y, scratch, s);
  if (fe != hipSuccess) return fe;
  const int64_t n = batch * k;
  mse_grad_kernel<<<blocks(n), kBlock, 0, s>>>(y, target, n, dy, partial);
  loss_final_kernel<<<1, 64, 0, s>>>(partial, static_cast<int>(blocks(n)), n, loss);
  dense_grad_weights_kernel<<<blocks(k * 65), kBlock, 0, s>>>(
      dy, k, h2, 64, batch, const_cast<float*>(g.w3), const_cast<float*>(g.b3));
  dense_grad_input_kernel<<<blocks(batch * 64), kBlock, 0, s>>>(dy, k, w.w3, h2, 64, batch, dh2);
  dense_grad_weights_kernel<<<blocks(64 * 65), kBlock, 0, s>>>(
      dh2, 64, h1, 64, batch, const_cast<float*>(g.w2), const_cast<float*>(g.b2));
  dense_grad_input_kernel<<<blocks(batch * 64), kBlock, 0, s>>>(dh2, 64, w.w2, h1, 64, batch, dh1);
  dense_grad_weights_kernel<<<blocks(64 * 17), kBlock, 0, s>>>(
      dh1, 64, x, 16, batch, const_cast<float*>(g.w1), const_cast<float*>(g.b1));
  return hipGetLastError();
}

hipError_t mlp_adam_launch(float* params, const float* grads, float* m, float* v,
                           int64_t n, float lr, float b1, float b2, float eps,
                           int64_t step, hipStream_t s) {
  const float bc1 = 1.0f - powf(b1, static_cast<float>(step));
  const float bc2 = 1.0f - powf(b2, static_cast<float>(step));
  adam_kernel<<<blocks(n), kBlock, 0, s>>>(params, grads, m, v, n, lr, b1, b2, eps, bc1, bc2);
  return hipGetLastError();
}

}  // namespace asc_hip

namespace asc_hip {
namespace {

__global__ __launch_bounds__(256) void normalize_kernel(const float* __restrict__ x,
                                                       const float* __restrict__ mean,
                                                       const float* __restrict__ stddev,
                                                       int64_t n, float* __restrict__ out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= n * 16) return;
  const int f = static_cast<int>(t % 16);
  const float sd = stddev[f] > 0.0f ? stddev[f] : 1.0f;
  out[t] = (x[t] - mean[f]) / sd;
}

}  // namespace

hipError_t normalize_launch(const float* x, const float* mean, const float* stddev,
                            int64_t n, float* out, hipStream_t s) {
  normalize_kernel<<<static_cast<unsigned>((n * 16 + 255) / 256), 256, 0, s>>>(x, mean, stddev, n, out);
  return hipGetLastError();
}

}  // namespace asc_hip
