/* C ABI of the warm-start network kernels (libvlbm.so, HIP, gfx950).
 *
 * No counterpart in the reference: it steps every case from rest.  These
 * entry points train and evaluate the small convolutional neural operator of
 * the learned warm start (curved_lbm/surrogate): six 3x3x3 convolutions of
 * 32 channels on a coarse lattice, BF16 operands on the MFMA units
 * (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.  Its prediction is
 * handed to the solver through the reference's initialize_from_numpy
 * (curved_boundary_warp3d.py:4216-4236).
 *
 * Layout.  A grid {batch, nx, ny, nz} of cells is stored with a one-cell
 * zero halo on every face: sample b, cell (x, y, z) sits at voxel
 * b * (nx+2)(ny+2)(nz+2) + ((x+1)(ny+2) + (y+1))(nz+2) + (z+1).  Activations
 * are channels-last, 32 BF16 values (64 bytes) per voxel; the last layer's
 * output is 32 FP32 values per voxel.  Kernels write interior voxels only,
 * so buffers that start zeroed keep a zero halo.  BF16 values are stored
 * as uint16_t bit patterns (round to nearest even on conversion).
 * Weights of a layer: BF16 [27 taps][32 out][32 in], tap = (dx+1)*9 +
 * (dy+1)*3 + (dz+1) for the neighbour offset (dx, dy, dz) in {-1, 0, 1}^3;
 * bias FP32 [32].
 *
 * Every entry point enqueues on `stream` (a hipStream_t, NULL for the null
 * stream), returns 0 on success or a negative code with vlbm_last_error().
 */
#ifndef VLBM_NN_H
#define VLBM_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLBM_NN_CH 32
#define VLBM_NN_TAPS 27
#define VLBM_NN_WEIGHTS (VLBM_NN_TAPS * VLBM_NN_CH * VLBM_NN_CH) /* per layer */
#define VLBM_NN_TARGET_CH 16 /* target channels per voxel; channel 15 is the voxel's loss weight */
#define VLBM_NN_WGRAD_PARTIAL (VLBM_NN_WEIGHTS + VLBM_NN_CH) /* floats per weight-gradient partial */

typedef struct {
    int32_t batch, nx, ny, nz;
} vlbm_nn_grid;

#define VLBM_NN_HIDDEN 0 /* z = conv + bias (BF16) and a = SiLU(z) (BF16) */
#define VLBM_NN_OUTPUT 1 /* y = conv + bias (FP32) */

/* One convolution layer, forward.  mode VLBM_NN_HIDDEN writes z_out and
 * a_out, VLBM_NN_OUTPUT writes y_out. */
int vlbm_nn_conv_forward(const vlbm_nn_grid *g, const uint16_t *in, const uint16_t *w, const float *bias,
                         int32_t mode, uint16_t *z_out, uint16_t *a_out, float *y_out, void *stream);
/* Gradient with respect to the layer input, times SiLU'(z_prev) of the
 * layer below: dz_prev = (sum_tap wt[tap] * dz[p + off(tap)]) * SiLU'(z_prev),
 * with wt the backward weights of vlbm_nn_pack_weights. */
int vlbm_nn_conv_backward_data(const vlbm_nn_grid *g, const uint16_t *dz, const uint16_t *wt,
                               const uint16_t *z_prev, uint16_t *dz_prev, void *stream);
/* Weight and bias gradients: dw[tap][o][i] = sum_p dz[p][o] in[p + off(tap)][i],
 * db[o] = sum_p dz[p][o], over all interior voxels of all samples (FP32).
 * partials: n_partials * VLBM_NN_WGRAD_PARTIAL floats of workspace;
 * n_partials in [1, 1024] sets the number of workgroups. */
int vlbm_nn_conv_backward_weight(const vlbm_nn_grid *g, const uint16_t *in, const uint16_t *dz, float *partials,
                                 int32_t n_partials, float *dw, float *db, void *stream);
/* Weighted squared error of the first n_out (<= 15) output channels:
 * loss = inv_norm * sum_p w_p sum_c chan_w[c] (y[p][c] - t[p][c])^2 with
 * t = target[p][c] and w_p = target[p][15]; writes dz = dloss/dy (BF16,
 * zero in channels >= n_out) and the loss (one double) to loss_dev.
 * partials: n_partials doubles of workspace, n_partials in [1, 4096]. */
int vlbm_nn_loss(const vlbm_nn_grid *g, const float *y, const float *target, const float *chan_w, int32_t n_out,
                 float inv_norm, uint16_t *dz, double *partials, int32_t n_partials, double *loss_dev, void *stream);
/* Rounds the FP32 master weights [27][32][32] of one layer to the BF16
 * forward weights w (same layout) and the backward-data weights
 * wt[tap][i][o] = w[26 - tap][o][i]. */
int vlbm_nn_pack_weights(const float *master, uint16_t *w, uint16_t *wt, void *stream);
/* Adam step over n parameters: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
 * p -= lr * (m * c1) / (sqrt(v * c2) + eps), with the bias corrections
 * c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t) computed by the caller. */
int vlbm_nn_adam(int32_t n, float *param, float *m, float *v, const float *grad, float lr, float b1, float b2,
                 float eps, float c1, float c2, void *stream);

#ifdef __cplusplus
}
#endif
#endif
