/* Small-MLP kernels for the learned CG warm start (libgrg_mlp.so), MI355X gfx950.
 *
 * No reference counterpart: the reference's mixed-model solver
 * (scripts/bolt_lmm_inf/core.py:1304-1353, bolt_conj_grad_solve) always starts
 * CG from zero. These kernels train and run a fully connected FP32 network
 * that maps a phenotype's top-r principal-component projections plus log(delta)
 * to the coefficients of a CG starting point.
 *
 * Layout: row-major, batch-major activations X[b * in + i]; weights W[o * in + i]
 * (out x in); biases b[o]. All pointers are device pointers owned by the caller;
 * nothing here allocates. Calls are asynchronous on the given hipStream_t and
 * return 0 on success, non-zero with grg_mlp_last_error() otherwise. Every
 * reduction runs in a fixed order, so results are deterministic.
 */
#ifndef GRG_MLP_H
#define GRG_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRG_MLP_ABI_VERSION 1

/* Y[b, o] = act(sum_i X[b, i] * W[o, i] + bias[o]); act = ReLU when relu != 0. */
int grg_mlp_linear_forward(const float* X, const float* W, const float* bias, float* Y, int32_t batch, int32_t in,
                           int32_t out, int32_t relu, void* stream);

/* Backward of grg_mlp_linear_forward given its output Y and the output
 * gradient dY: G = dY * (relu ? (Y > 0) : 1);
 * dW[o, i] = sum_b G[b, o] X[b, i]; dbias[o] = sum_b G[b, o];
 * dX[b, i] = sum_o G[b, o] W[o, i] (skipped when dX is NULL). */
int grg_mlp_linear_backward(const float* X, const float* W, const float* Y, const float* dY, float* dX, float* dW,
                            float* dbias, int32_t batch, int32_t in, int32_t out, int32_t relu, void* stream);

/* Weighted squared error: loss[0] = (1 / batch) sum_b sum_j w[j] (P - T)^2,
 * dP = (2 / batch) w[j] (P - T). w may be NULL (all ones). */
int grg_mlp_mse(const float* P, const float* T, const float* w, float* dP, float* loss, int32_t batch, int32_t n,
                void* stream);

/* One Adam step (Kingma & Ba) on n parameters with bias correction for step t >= 1:
 * m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2;
 * p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps). */
int grg_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2, float eps,
                 int32_t t, void* stream);

const char* grg_mlp_last_error(void);
int32_t grg_mlp_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GRG_MLP_H */
