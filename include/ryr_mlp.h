/*
 * ryr_mlp.h - C ABI of the beat surrogate: a small MLP that predicts, per
 * beat, the EAD probability and APD90 from the cell parameters and the
 * previous beat's summary, used to steer parameter sweeps of the CRU engine.
 *
 * The reference has no counterpart (it simulates every point of a sweep);
 * this is the learned part of the steering scheme and is not on the
 * reference's hot path.
 *
 * Architecture (fixed): 11 inputs -> 64 GELU -> 64 GELU -> 2 outputs
 * (EAD logit, APD90 in the caller's normalised units).  GELU is the tanh form
 * 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).  Matrix products run on
 * gfx950 MFMA with BF16 weights and activations and FP32 accumulation; the
 * master weights, Adam moments, biases, activation functions and losses are
 * FP32.
 *
 * Parameters are one packed FP32 vector of RYR_MLP_NPARAM values:
 *   W1 [32][64]  (rows 11..31 are padding and stay 0)   offset 0
 *   b1 [64]                                              offset 2048
 *   W2 [64][64]                                          offset 2112
 *   b2 [64]                                              offset 6208
 *   W3 [64][16]  (columns 2..15 are padding and stay 0) offset 6272
 *   b3 [16]      (entries 2..15 are padding)             offset 7296
 * Row-major: W[k][j] multiplies input k into unit j.
 *
 * Loss over n samples with weights s_i (1 for a plain sample):
 *   L = sum_i s_i [ BCEwithLogits(z_i0, y_i) + w_apd m_i (z_i1 - t_i)^2 ] / sum_i s_i
 * with y_i in {0,1} the EAD label, t_i the APD90 target and m_i in {0,1}
 * whether the beat has an APD90.
 *
 * All calls return 0 or a negative RYR_E_* code (ryr_cru.h); host pointers
 * and sizes only.
 */
#ifndef RYR_MLP_H
#define RYR_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RYR_MLP_NIN    11
#define RYR_MLP_NHID   64
#define RYR_MLP_NOUT   2
#define RYR_MLP_NPARAM 7312

typedef struct ryr_mlp ryr_mlp;

/* A network on `device` with all parameters 0 (set them before use). */
int ryr_mlp_create(int device, ryr_mlp **out);
void ryr_mlp_destroy(ryr_mlp *m);

/* Copy the packed FP32 parameters in / out (n = RYR_MLP_NPARAM); setting
 * them also resets the Adam state and its step count. */
int ryr_mlp_set_params(ryr_mlp *m, const float *params, int64_t n);
int ryr_mlp_get_params(ryr_mlp *m, float *params, int64_t n);

/* Inference: x is n x 11 (row-major); out is n x 2: sigmoid(z0) (EAD
 * probability) and z1 (APD90, normalised). */
int ryr_mlp_forward(ryr_mlp *m, const float *x, int64_t n, float *out);

/* Loss and its gradient with respect to the packed parameters (grad: n
 * RYR_MLP_NPARAM values) at the current parameters, without updating them.
 * y: EAD labels, t: APD90 targets, mask: 1 where t is valid, s: sample
 * weights (null = all 1). */
int ryr_mlp_grad(ryr_mlp *m, const float *x, const float *y, const float *t, const float *mask,
                 const float *s, int64_t n, float w_apd, float *loss, float *grad);

/* Adam training: `steps` updates on minibatches of `batch` samples drawn in
 * order from a permutation of the n samples (a new permutation, from `seed`,
 * each pass).  loss_trace (may be null) receives the minibatch loss of every
 * step. */
int ryr_mlp_train(ryr_mlp *m, const float *x, const float *y, const float *t, const float *mask,
                  int64_t n, float w_apd, int64_t steps, int64_t batch, float lr, float beta1,
                  float beta2, float eps, uint64_t seed, float *loss_trace);

const char *ryr_mlp_last_error(const ryr_mlp *m);

#ifdef __cplusplus
}
#endif

#endif
