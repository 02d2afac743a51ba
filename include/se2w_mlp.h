/*
 * se2w_mlp.h -- C ABI of the seismogram surrogate in libse2w_hip.so: a
 * multilayer perceptron (GELU on hidden layers) trained and evaluated on one
 * MI355X by hand-written kernels: BF16 matrix-core (MFMA) products with FP32
 * accumulation for the forward pass and both backward products, FP32 master
 * weights and an Adam update that refreshes the BF16 copies.
 *
 * The reference has no counterpart (it only runs the solver); the surrogate
 * maps a source position and a coarse velocity-perturbation grid to receiver
 * seismograms, and the solver (se2w.h) both generates its training data and
 * measures its error.
 *
 * Shapes: widths[0] inputs (any size; padded to a multiple of 32 inside),
 * hidden and output widths multiples of 64, batch a multiple of 64.
 * Parameters are exchanged as one FP32 array: for each layer l, W_l
 * [widths[l+1]][widths[l]] row-major, then b_l [widths[l+1]].
 * Status codes and messages as in se2w.h.
 */
#ifndef SE2W_MLP_H
#define SE2W_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct se2w_mlp se2w_mlp;

/* nlayers linear layers, widths[0..nlayers]; weights ~ U(-1/sqrt(fan_in),
 * 1/sqrt(fan_in)) and biases likewise, drawn from a splitmix64 stream. */
int se2w_mlp_create(int32_t nlayers, const int32_t *widths, int32_t batch, uint64_t seed, int32_t device,
                    se2w_mlp **mlp);
int se2w_mlp_param_count(se2w_mlp *mlp, int64_t *count);
int se2w_mlp_get_params(se2w_mlp *mlp, float *params);
/* Sets the FP32 master weights (and their BF16 copies); Adam moments are reset. */
int se2w_mlp_set_params(se2w_mlp *mlp, const float *params);
/* y[n][out] = network(x[n][in]), host arrays, processed in batches. */
int se2w_mlp_forward(se2w_mlp *mlp, int32_t n, const float *x, float *y);
/* Mean-squared-error loss of one batch (n == batch) and its gradient with
 * respect to every parameter (same layout as the parameters). */
int se2w_mlp_gradient(se2w_mlp *mlp, int32_t n, const float *x, const float *target, float *loss, float *grad);
/* Training set, kept in device memory: x[n][in], target[n][out]. */
int se2w_mlp_set_data(se2w_mlp *mlp, int32_t n, const float *x, const float *target);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over `epochs` passes of shuffled
 * full batches (a remainder smaller than a batch is skipped); epoch_loss
 * [epochs] receives each epoch's mean batch loss (may be NULL). */
int se2w_mlp_train(se2w_mlp *mlp, int32_t epochs, float lr, uint64_t shuffle_seed, float *epoch_loss);
/* Learning rate of se2w_mlp_train: cosine = 0 (default) keeps lr constant;
 * cosine = 1 decays it as lr (1 + cos(pi s / S)) / 2 over the S batch steps
 * of the call (Loshchilov & Hutter 2017, without restarts). */
int se2w_mlp_set_lr_schedule(se2w_mlp *mlp, int32_t cosine);
/* Decoupled weight decay of the Adam update (AdamW, Loshchilov & Hutter
 * 2019): w -= lr (mhat / (sqrt(vhat) + eps) + wd w).  Default 0 (Adam). */
int se2w_mlp_set_weight_decay(se2w_mlp *mlp, float wd);
/* One Adam step with the given gradient (same layout as the parameters). */
int se2w_mlp_adam_step(se2w_mlp *mlp, const float *grad, float lr);
/* Wall time of the last se2w_mlp_train and of the last se2w_mlp_forward's
 * device work (HIP events), in ms. */
int se2w_mlp_get_timing(se2w_mlp *mlp, double *train_ms, double *forward_ms);
int se2w_mlp_destroy(se2w_mlp *mlp);

#ifdef __cplusplus
}
#endif

#endif
