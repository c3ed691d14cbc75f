/* igap_ws.h -- C ABI of the learned Newton warm start (no counterpart in the
 * reference: IGAP4 starts each load or time step from the previous solution,
 * src/interface.c:252-259,319).
 *
 * A small 3-D convolutional network on the control-point grid (4 layers of
 * 3x3x3 convolutions, 16 channels, per-layer dilation, ReLU between layers)
 * maps features of the step history to a correction of the linear
 * extrapolation of the next solution.  Inference and training run on the
 * GPU's BF16 matrix cores with FP32 accumulation; the master parameters and
 * Adam moments are FP32.  The caller (igap4_amd/warmstart.py) builds the
 * features, applies the correction and accepts the guess only if its residual
 * norm beats the extrapolation's.
 *
 * Arrays are host arrays, channels-last over the grid: x[v][cin], y and
 * target [v][cout] with v = (i*ny + j)*nz + k (the assembly's control-point
 * numbering).  cin, cout <= 16.  Parameters (FP32) are laid out per layer as
 * W[27][16 out][16 in] (tap t = 9(dx+1) + 3(dy+1) + (dz+1)) followed by b[16];
 * padded channels stay zero.  Every call is synchronous and returns an
 * IGAP_* status; igap_gpu_last_error() gives the message.
 */
#ifndef IGAP_WS_H
#define IGAP_WS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IGAP_WS_LAYERS 4
#define IGAP_WS_CHANNELS 16

typedef struct igap_ws_ctx igap_ws_ctx;

/* grid = control points per direction; dil = dilation of each layer; the
 * hidden layers are He-uniform initialised from seed, the last layer is zero
 * (the untrained network returns the extrapolation unchanged). */
int igap_ws_create(int device, const int grid[3], int cin, int cout, const int dil[IGAP_WS_LAYERS], uint64_t seed,
                   igap_ws_ctx **out);
int igap_ws_destroy(igap_ws_ctx *ctx);
int igap_ws_nparams(igap_ws_ctx *ctx, int64_t *n);
int igap_ws_get_params(igap_ws_ctx *ctx, float *params);
int igap_ws_set_params(igap_ws_ctx *ctx, const float *params);
/* y = net(x) */
int igap_ws_forward(igap_ws_ctx *ctx, const float *x, float *y);
/* loss = mean over v and c < cout of (net(x) - target)^2 and its gradient
 * with respect to every parameter (params layout), without an update */
int igap_ws_gradients(igap_ws_ctx *ctx, const float *x, const float *target, double *loss, float *grad);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the loss above; loss
 * is the value before the update */
int igap_ws_train_step(igap_ws_ctx *ctx, const float *x, const float *target, float lr, double *loss);
/* one 16x16x16 BF16 product of A[i][k] = 1 + i + k/4, B[k][n] = 2 - n/2 + k
 * (lane l supplies A[l&15][4(l>>4)+j] and B[4(l>>4)+j][l&15]); out receives
 * the 4 x 64 raw FP32 result registers (out[r*64 + l]) */
int igap_ws_selftest_mfma_bf16(float *out);

#ifdef __cplusplus
}
#endif
#endif
