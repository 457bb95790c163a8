/*
 * mpk_mlp.h -- C ABI of the learned reluctivity predictor's network.
 *
 * No reference counterpart: MagPackage has no learned component.  The
 * network maps per-element features (previous-step B, its extrapolation,
 * region) to a normalised log-reluctivity for the next Newton step's warm
 * start.  Architecture: MPK_MLP_IN inputs -> 3 hidden layers of
 * MPK_MLP_HIDDEN tanh units -> 1 linear output.  Hidden-layer weights are
 * used as BF16 on the matrix cores (v_mfma_f32_16x16x16bf16_1k, FP32
 * accumulation); biases and the output layer stay FP32; the optimiser keeps
 * FP32 master parameters.
 *
 * Parameter vector (FP32, MPK_MLP_NPARAM entries, this order):
 *   W1 [64 x 16] (out-major), b1 [64], W2 [64 x 64], b2 [64],
 *   W3 [64 x 64], b3 [64], w4 [64], b4 [1].
 * BF16 copy (MPK_MLP_NBF16 entries): W1, W2, W3 in the same order.
 * All arrays are device arrays; calls are asynchronous on `stream`.
 * Status codes are those of mpk_assembly.h.
 */
#ifndef MPK_MLP_H
#define MPK_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPK_MLP_IN 16
#define MPK_MLP_HIDDEN 64
#define MPK_MLP_NPARAM 9473
#define MPK_MLP_NBF16 9216

/* BF16 copies of the three hidden-layer weight matrices (round to nearest even). */
int mpk_mlp_sync_bf16(const float* params, uint16_t* w_bf16, void* stream);

/* y[n] = net(X[n x MPK_MLP_IN]) (X row-major FP32, rounded to BF16 on load). */
int mpk_mlp_forward(const float* params, const uint16_t* w_bf16, const float* X, int64_t n,
                    float* y, void* stream);

/* Floats of device workspace mpk_mlp_loss_grad needs for n rows. */
int mpk_mlp_work_size(int64_t n, int64_t* size_out);

/* Mean squared error L = sum_r (y_r - t_r)^2 / n and its gradient with
 * respect to every parameter: grad [MPK_MLP_NPARAM], loss [1] (device).
 * Partial sums per wave are reduced in a fixed order: deterministic. */
int mpk_mlp_loss_grad(const float* params, const uint16_t* w_bf16, const float* X,
                      const float* t, int64_t n, float* grad, float* loss, float* work,
                      void* stream);

/* One Adam step (bias-corrected, step >= 1) on the FP32 master parameters
 * with moment buffers m, v [MPK_MLP_NPARAM]; refreshes the BF16 copies. */
int mpk_mlp_adam(float* params, float* m, float* v, const float* grad, int64_t step, float lr,
                 float beta1, float beta2, float eps, uint16_t* w_bf16, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPK_MLP_H */
