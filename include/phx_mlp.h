/*
 * phx_mlp.h -- learned con2prim bracket predictor (libphx.so).  No counterpart in the
 * reference: phoebus's robust con2prim (con2prim_robust.hpp:374-375) brackets the root
 * mu in [0, 1] and splits it at the previous stage's mu.  This network predicts a narrow
 * bracket from scale-free invariants of the conserved state; the reference's regula falsi
 * then runs to the same tolerance inside it, or on [0, 1] when the residual does not
 * change sign across it, so only the iteration count changes.
 *
 * Network: 4 -> 64 -> 64 -> 64 -> 2, ReLU, BF16 MFMA operands with FP32 accumulation,
 * PHX_MLP_NPARAM = 8770 FP32 master parameters in the order W1[64][4] b1[64] W2[64][64]
 * b2[64] W3[64][64] b3[64] W4[2][64] b4[2] (row = output neuron).
 * Features: (log(q + 1e-12), log(r^2 + 1e-12), log(b^2 + 1e-12), log((r.b)^2 + 1e-24))
 * standardized by (x - mean) / std, with q = tau/D, r_i = S_i/D, b^i = B^i/sqrt(D) as in
 * con2prim_robust.hpp:327-361.  Outputs are relative to the prior z0 = max(1 + q - b^2/2, 1)
 * (an estimate of hW = 1/mu from the energy equation, between hW/Gamma and hW for b = 0,
 * evaluated in fp32 from the standardized features): o1 = log(c z0) for the bracket centre c, o2 = log of the
 * half-width relative to c; the bracket is c (1 -+ max(width_scale * exp(o2), width_min))
 * clipped to [0, 1].
 * All array arguments are device pointers.
 */
#ifndef PHX_MLP_H_
#define PHX_MLP_H_

#include <stdint.h>

#include "phx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_MLP_NPARAM 8770

typedef struct phx_mlp phx_mlp;

int phx_mlp_create(phx_ctx *ctx, phx_mlp **mlp, uint64_t seed);
int phx_mlp_destroy(phx_mlp *mlp);
int phx_mlp_set_params(phx_ctx *ctx, phx_mlp *mlp, const float *host_params);
int phx_mlp_get_params(phx_ctx *ctx, const phx_mlp *mlp, float *host_params);
int phx_mlp_set_normalization(phx_ctx *ctx, phx_mlp *mlp, const float mean[4],
                              const float std_[4], float width_scale, float width_min);

/* standardized features [4][n] (float) of n batched conserved states (layout of
 * phx_con2prim_batch) */
int phx_con2prim_features(phx_ctx *ctx, const phx_fluid_params *p, const phx_mlp *mlp,
                          int64_t n, const double *coords, const double *cons,
                          float *features, phx_stream stream);

/* network outputs [2][n] (o1, o2) for features [4][n] */
int phx_mlp_forward(phx_ctx *ctx, const phx_mlp *mlp, int64_t n, const float *features,
                    float *out, phx_stream stream);

/* Training.  Loss per sample (averaged over n), with t = log(mu* z0) and mu* the plain
 * solver's root: (o1 - t)^2 + lam (o2 - log(|1 - exp(t - o1)| + 1e-12))^2, the second
 * target held fixed.
 * phx_mlp_gradient accumulates the loss gradient into the context-free gradient buffer
 * of the network (and the loss into *loss_out, device, nullable); phx_mlp_adam_step
 * applies Adam (beta 0.9, 0.999) and clears the gradient; phx_mlp_train_step does both. */
int phx_mlp_gradient(phx_ctx *ctx, phx_mlp *mlp, int64_t n, const float *features,
                     const float *target_mu, float lam, float *loss_out, phx_stream stream);
int phx_mlp_read_gradient(phx_ctx *ctx, phx_mlp *mlp, float *host_grad);
int phx_mlp_zero_gradient(phx_ctx *ctx, phx_mlp *mlp);
int phx_mlp_adam_step(phx_ctx *ctx, phx_mlp *mlp, float lr, phx_stream stream);
int phx_mlp_train_step(phx_ctx *ctx, phx_mlp *mlp, int64_t n, const float *features,
                       const float *target_mu, float lr, float lam, float *loss_out,
                       phx_stream stream);

/* phx_con2prim_batch with the learned bracket.  evals (nullable) = residual evaluations
 * including the bracket ends; fallback (nullable) = 1 where the bracket missed the root
 * and the plain [0, 1] solve ran (iters then counts that solve). */
int phx_con2prim_batch_learned(phx_ctx *ctx, const phx_fluid_params *p, const phx_mlp *mlp,
                               int64_t n, const double *coords, double *cons, double *prim,
                               double *mu, double *fail, int32_t *iters, int32_t *evals,
                               int32_t *fallback, phx_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PHX_MLP_H_ */
