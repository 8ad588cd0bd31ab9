/*
 * mcls_mlp.h -- the importance network of the learned proposal.
 *
 * No counterpart in the reference: MCLS samples transitions only from
 * P_ij = |h_ij| / W_i (src/MCLS_AlmostOptimalDomain_impl.hpp:342-363).  The
 * build adds a small network that predicts a per-state importance g from
 * local features; the walk then samples P_ij proportional to |h_ij| g_j with
 * unbiased weights (mcls_gpu_set_importance, mcls_gpu_set_importance_model).
 *
 * Network: 8 inputs -> 64 -> 64 -> 64 -> 1, ReLU hidden layers.  Matrix
 * products run on MFMA with bfloat16 operands and FP32 accumulation; the
 * master parameters are FP32 in the flat layout
 *   W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4[1]
 * (z = a W + b).  The importance is g = exp(clip(o, -4, 4)).
 * Loss: mean over rows of (o - t)^2; training uses Adam.
 */
#ifndef MCLS_MLP_H
#define MCLS_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCLS_MLP_INPUTS 8
#define MCLS_MLP_HIDDEN 64
#define MCLS_MLP_PARAMS 8961

typedef struct mcls_mlp mcls_mlp;

/* Network on `device` with He-normal weights drawn from `seed`. */
int mcls_mlp_create(mcls_mlp** m, int device, uint64_t seed);
int mcls_mlp_destroy(mcls_mlp* m);
int mcls_mlp_get_params(mcls_mlp* m, float* params);        /* host, MCLS_MLP_PARAMS */
int mcls_mlp_set_params(mcls_mlp* m, const float* params);  /* host, MCLS_MLP_PARAMS */
/* out[r] = o for rows x 8 inputs X (FP32).  ptr_kind as in mcls_gpu.h. */
int mcls_mlp_forward(mcls_mlp* m, const float* X, int64_t rows, float* out, int ptr_kind);
/* Mean squared error against T and its gradient (MCLS_MLP_PARAMS, host). */
int mcls_mlp_loss_grad(mcls_mlp* m, const float* X, const float* T, int64_t rows,
                       double* loss, float* grad, int ptr_kind);
/* Device-resident training set: rows of (X, T) appended, then trained on. */
int mcls_mlp_dataset_clear(mcls_mlp* m);
int mcls_mlp_dataset_append(mcls_mlp* m, const float* X, const float* T, int64_t rows, int ptr_kind);
int mcls_mlp_dataset_rows(mcls_mlp* m, int64_t* rows);
/* Adam over the stored set: `epochs` passes of mini-batches of batch_rows
 * contiguous rows in a seeded random order; final_loss = mean loss of the
 * last epoch. */
int mcls_mlp_train(mcls_mlp* m, int epochs, int64_t batch_rows, float lr, uint64_t seed,
                   double* final_loss);

#ifdef __cplusplus
}
#endif
#endif
