// Small fully connected FP32 network kernels (forward, backward, loss, Adam)
// for the learned CG warm start, gfx950. See include/grg_mlp.h.
//
// The network is tiny (tens of inputs, 256-wide hidden layers, batches of a
// few thousand), so each product is a 16 x 16 output tile per 256-thread
// workgroup with both operand tiles staged through LDS; every output element
// is one thread's sum over the inner dimension in ascending order, so results
// do not depend on the launch (no atomics, no split reductions).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/grg_mlp.h"

namespace {

thread_local std::string g_last_error;

int fail(const std::string& msg, int code = 1) {
  g_last_error = msg;
  return code;
}

int launched() {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
  return 0;
}

constexpr int kTile = 16;

// Y[b, o] = act(sum_i X[b, i] W[o, i] + bias[o])
__global__ __launch_bounds__(256) void linear_forward_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                                             const float* __restrict__ bias, float* __restrict__ Y,
                                                             int batch, int in, int out, int relu) {
  __shared__ float xs[kTile][kTile + 1];
  __shared__ float ws[kTile][kTile + 1];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int b = blockIdx.y * kTile + ty;  // row of Y
  const int o = blockIdx.x * kTile + tx;  // column of Y
  float acc = 0.f;
  for (int i0 = 0; i0 < in; i0 += kTile) {
    const int bx = blockIdx.y * kTile + ty, ix = i0 + tx;
    xs[ty][tx] = (bx < batch && ix < in) ? X[(int64_t)bx * in + ix] : 0.f;
    const int ow = blockIdx.x * kTile + ty;
    ws[ty][tx] = (ow < out && ix < in) ? W[(int64_t)ow * in + ix] : 0.f;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kTile; ++t) acc += xs[ty][t] * ws[tx][t];
    __syncthreads();
  }
  if (b < batch && o < out) {
    float y = acc + (bias ? bias[o] : 0.f);
    if (relu && y < 0.f) y = 0.f;
    Y[(int64_t)b * out + o] = y;
  }
}

__device__ __forceinline__ float grad_at(const float* __restrict__ Y, const float* __restrict__ dY, int64_t idx, int relu) {
  const float g = dY[idx];
  return (relu && !(Y[idx] > 0.f)) ? 0.f : g;
}

// dW[o, i] = sum_b G[b, o] X[b, i]
__global__ __launch_bounds__(256) void linear_dw_kernel(const float* __restrict__ X, const float* __restrict__ Y,
                                                        const float* __restrict__ dY, float* __restrict__ dW, int batch,
                                                        int in, int out, int relu) {
  __shared__ float gs[kTile][kTile + 1];  // [b][o]
  __shared__ float xs[kTile][kTile + 1];  // [b][i]
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int o = blockIdx.y * kTile + ty;
  const int i = blockIdx.x * kTile + tx;
  float acc = 0.f;
  for (int b0 = 0; b0 < batch; b0 += kTile) {
    const int bl = b0 + ty;
    const int og = blockIdx.y * kTile + tx;
    gs[ty][tx] = (bl < batch && og < out) ? grad_at(Y, dY, (int64_t)bl * out + og, relu) : 0.f;
    xs[ty][tx] = (bl < batch && i < in) ? X[(int64_t)bl * in + i] : 0.f;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kTile; ++t) acc += gs[t][ty] * xs[t][tx];
    __syncthreads();
  }
  if (o < out && i < in) dW[(int64_t)o * in + i] = acc;
}

// dbias[o] = sum_b G[b, o]
__global__ __launch_bounds__(256) void linear_db_kernel(const float* __restrict__ Y, const float* __restrict__ dY,
                                                        float* __restrict__ db, int batch, int out, int relu) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= out) return;
  float acc = 0.f;
  for (int b = 0; b < batch; ++b) acc += grad_at(Y, dY, (int64_t)b * out + o, relu);
  db[o] = acc;
}

// dX[b, i] = sum_o G[b, o] W[o, i]
__global__ __launch_bounds__(256) void linear_dx_kernel(const float* __restrict__ W, const float* __restrict__ Y,
                                                        const float* __restrict__ dY, float* __restrict__ dX, int batch,
                                                        int in, int out, int relu) {
  __shared__ float gs[kTile][kTile + 1];  // [b][o]
  __shared__ float ws[kTile][kTile + 1];  // [o][i]
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int b = blockIdx.y * kTile + ty;
  const int i = blockIdx.x * kTile + tx;
  float acc = 0.f;
  for (int o0 = 0; o0 < out; o0 += kTile) {
    const int og = o0 + tx;
    gs[ty][tx] = (b < batch && og < out) ? grad_at(Y, dY, (int64_t)b * out + og, relu) : 0.f;
    const int ow = o0 + ty;
    ws[ty][tx] = (ow < out && i < in) ? W[(int64_t)ow * in + i] : 0.f;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kTile; ++t) acc += gs[ty][t] * ws[t][tx];
    __syncthreads();
  }
  if (b < batch && i < in) dX[(int64_t)b * in + i] = acc;
}

// One workgroup: dP over all elements and the loss by a fixed-order tree reduction.
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ P, const float* __restrict__ T,
                                                  const float* __restrict__ w, float* __restrict__ dP,
                                                  float* __restrict__ loss, int batch, int n) {
  __shared__ double part[256];
  const int64_t total = (int64_t)batch * n;
  const float inv = 1.f / (float)batch;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int j = (int)(e % n);
    const float wj = w ? w[j] : 1.f;
    const float d = P[e] - T[e];
    dP[e] = 2.f * inv * wj * d;
    acc += (double)wj * (double)d * (double)d;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = (float)(part[0] / (double)batch);
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int check_dims(int32_t batch, int32_t in, int32_t out) {
  if (batch < 1 || in < 1 || out < 1) return fail("batch, in and out must be >= 1");
  if ((int64_t)batch * in > 0x7fffffffLL || (int64_t)batch * out > 0x7fffffffLL || (int64_t)in * out > 0x7fffffffLL)
    return fail("dimensions too large");
  return 0;
}

}  // namespace

extern "C" {

int grg_mlp_linear_forward(const float* X, const float* W, const float* bias, float* Y, int32_t batch, int32_t in,
                           int32_t out, int32_t relu, void* stream) {
  g_last_error.clear();
  if (!X || !W || !Y) return fail("null pointer");
  if (int rc = check_dims(batch, in, out)) return rc;
  dim3 grid((out + kTile - 1) / kTile, (batch + kTile - 1) / kTile);
  hipLaunchKernelGGL(linear_forward_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), X, W, bias, Y, batch,
                     in, out, relu);
  return launched();
}

int grg_mlp_linear_backward(const float* X, const float* W, const float* Y, const float* dY, float* dX, float* dW,
                            float* dbias, int32_t batch, int32_t in, int32_t out, int32_t relu, void* stream) {
  g_last_error.clear();
  if (!X || !W || !Y || !dY || !dW || !dbias) return fail("null pointer");
  if (int rc = check_dims(batch, in, out)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(linear_dw_kernel, dim3((in + kTile - 1) / kTile, (out + kTile - 1) / kTile), dim3(256), 0, s, X, Y,
                     dY, dW, batch, in, out, relu);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(linear_db_kernel, dim3((out + 255) / 256), dim3(256), 0, s, Y, dY, dbias, batch, out, relu);
  if (int rc = launched()) return rc;
  if (dX) {
    hipLaunchKernelGGL(linear_dx_kernel, dim3((in + kTile - 1) / kTile, (batch + kTile - 1) / kTile), dim3(256), 0, s, W,
                       Y, dY, dX, batch, in, out, relu);
    if (int rc = launched()) return rc;
  }
  return 0;
}

int grg_mlp_mse(const float* P, const float* T, const float* w, float* dP, float* loss, int32_t batch, int32_t n,
                void* stream) {
  g_last_error.clear();
  if (!P || !T || !dP || !loss) return fail("null pointer");
  if (batch < 1 || n < 1) return fail("batch and n must be >= 1");
  hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(256), 0, static_cast<hipStream_t>(stream), P, T, w, dP, loss, batch, n);
  return launched();
}

int grg_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps,
                 int32_t t, void* stream) {
  g_last_error.clear();
  if (!p || !g || !m || !v) return fail("null pointer");
  if (n < 0) return fail("n < 0");
  if (t < 1) return fail("Adam step t must be >= 1");
  if (n == 0) return 0;
  const float c1 = 1.f - std::pow(beta1, (float)t);
  const float c2 = 1.f - std::pow(beta2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), p, g,
                     m, v, n, lr, beta1, beta2, eps, c1, c2);
  return launched();
}

const char* grg_mlp_last_error(void) { return g_last_error.c_str(); }

int32_t grg_mlp_abi_version(void) { return GRG_MLP_ABI_VERSION; }

}  // extern "C"
