// Small fully connected FP32 network kernels (forward, backward, loss, Adam)
// for the learned CG warm start, gfx950. See include/grg_mlp.h.
//
// The three products of a linear layer (forward, weight gradient, input
// gradient) run on the matrix cores: each wave owns one 16 x 16 output tile and
// steps through the inner dimension four at a time with v_mfma_f32_16x16x4f32
// (FP32 products and sums), four waves per 32 x 32 block of the output. The
// network is small (tens of inputs, 256-wide hidden layers, batches of a few
// thousand), so operands are read straight from global memory (they stay in
// L2). Every output element is one wave's sum in a fixed order, so results do
// not depend on the launch (no atomics, no split reductions).

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

typedef float MfmaAcc __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float grad_at(const float* __restrict__ Y, const float* __restrict__ dY, int64_t idx, int relu) {
  const float g = dY[idx];
  return (relu && !(Y[idx] > 0.f)) ? 0.f : g;
}

enum LinearProduct { kForward = 0, kWeightGrad = 1, kInputGrad = 2 };

// C = A B on the matrix cores, with G = dY masked by relu'(Y):
//   kForward:    Y[b, o]  = act(sum_i X[b, i] W[o, i] + bias[o])  (M = batch, N = out, K = in)
//   kWeightGrad: dW[o, i] = sum_b G[b, o] X[b, i]                 (M = out, N = in, K = batch)
//   kInputGrad:  dX[b, i] = sum_o G[b, o] W[o, i]                 (M = batch, N = in, K = out)
// Wave w of the block owns the 16 x 16 tile at rows m0 = (2 blockIdx.y + w / 2) * 16,
// columns n0 = (2 blockIdx.x + w % 2) * 16. Per step of four k, lane r + 16 q supplies
// A[m0 + r][k0 + q] and B[k0 + q][n0 + r] (zero past the matrix edges), and after the
// last step holds C[m0 + 4 q + t][n0 + r] in accumulator t.
template <int P>
__global__ __launch_bounds__(256) void linear_mfma_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                                          const float* __restrict__ Y, const float* __restrict__ dY,
                                                          const float* __restrict__ bias, float* __restrict__ C,
                                                          int batch, int in, int out, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = P == kWeightGrad ? out : batch;
  const int N = P == kForward ? out : in;
  const int K = P == kForward ? in : (P == kWeightGrad ? batch : out);
  const int m0 = ((int)blockIdx.y * 2 + (wave >> 1)) * 16;
  const int n0 = ((int)blockIdx.x * 2 + (wave & 1)) * 16;
  const int r = lane & 15, q = lane >> 4;
  const int m = m0 + r, n = n0 + r;
  MfmaAcc acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + q;
    float a = 0.f, bv = 0.f;
    if (k < K) {
      if (P == kForward) {
        if (m < M) a = X[(int64_t)m * in + k];
        if (n < N) bv = W[(int64_t)n * in + k];
      } else if (P == kWeightGrad) {
        if (m < M) a = grad_at(Y, dY, (int64_t)k * out + m, relu);
        if (n < N) bv = X[(int64_t)k * in + n];
      } else {
        if (m < M) a = grad_at(Y, dY, (int64_t)m * out + k, relu);
        if (n < N) bv = W[(int64_t)k * in + n];
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc, 0, 0, 0);
  }
  if (n >= N) return;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int mt = m0 + 4 * q + t;
    if (mt >= M) continue;
    float y = acc[t];
    if (P == kForward) {
      if (bias) y += bias[n];
      if (relu && y < 0.f) y = 0.f;
    }
    C[(int64_t)mt * N + n] = y;
  }
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
  dim3 grid((out + 31) / 32, (batch + 31) / 32);
  hipLaunchKernelGGL(linear_mfma_kernel<kForward>, grid, dim3(256), 0, static_cast<hipStream_t>(stream), X, W,
                     static_cast<const float*>(nullptr), static_cast<const float*>(nullptr), bias, Y, batch, in, out, relu);
  return launched();
}

int grg_mlp_linear_backward(const float* X, const float* W, const float* Y, const float* dY, float* dX, float* dW,
                            float* dbias, int32_t batch, int32_t in, int32_t out, int32_t relu, void* stream) {
  g_last_error.clear();
  if (!X || !W || !Y || !dY || !dW || !dbias) return fail("null pointer");
  if (int rc = check_dims(batch, in, out)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(linear_mfma_kernel<kWeightGrad>, dim3((in + 31) / 32, (out + 31) / 32), dim3(256), 0, s, X,
                     static_cast<const float*>(nullptr), Y, dY, static_cast<const float*>(nullptr), dW, batch, in, out, relu);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(linear_db_kernel, dim3((out + 255) / 256), dim3(256), 0, s, Y, dY, dbias, batch, out, relu);
  if (int rc = launched()) return rc;
  if (dX) {
    hipLaunchKernelGGL(linear_mfma_kernel<kInputGrad>, dim3((in + 31) / 32, (batch + 31) / 32), dim3(256), 0, s,
                       static_cast<const float*>(nullptr), W, Y, dY, static_cast<const float*>(nullptr), dX, batch, in, out,
                       relu);
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
