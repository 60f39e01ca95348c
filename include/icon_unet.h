/* icon_unet.h -- C ABI of the learned warm start.
 *
 * No counterpart in the reference: a 2D U-Net (4 levels, 32/64/128/256
 * channels, BF16 weights on MFMA with FP32 accumulation, FP32 master weights
 * and Adam) that maps early-iteration ICON slices to converged ones.  Its
 * output seeds icon_reconstruct's m_init; plain ICON (m_init = NULL) stays the
 * fallback.  Images are [B][H][W] float32 device buffers; H and W multiples
 * of 8.  Return 0 on success, non-zero with icon_unet_last_error() otherwise.
 */
#ifndef ICON_UNET_H
#define ICON_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icon_unet icon_unet;

const char* icon_unet_last_error(void);
/* A network for batches of up to max_batch images of H x W on `device`;
 * train != 0 also allocates gradients and optimizer state. */
int icon_unet_create(icon_unet** net, int device, int max_batch, int H, int W, int train);
int icon_unet_destroy(icon_unet* net);
int icon_unet_param_count(const icon_unet* net, int64_t* count);
/* He-normal initial weights from a seed. */
int icon_unet_init(icon_unet* net, uint64_t seed);
/* Flat FP32 parameter vector: per 3x3 convolution W[9][c_out][c_in] then
 * bias[c_out] (14 convolutions in network order), then the 1x1 head w[32], b. */
int icon_unet_get_params(const icon_unet* net, float* host, int64_t count);
int icon_unet_set_params(icon_unet* net, const float* host, int64_t count);
/* out = U-Net(in), B images. */
int icon_unet_forward(icon_unet* net, int B, const float* d_in, float* d_out, void* stream);
/* Learned warm start of S slices [S][H][W] (any S; batches of max_batch):
 *   d_out = d_m + s * U-Net(d_m / s),  s = RMS of the slice (1 if zero),
 * i.e. the network predicts the scaled change from an early ICON iterate to
 * a converged one.  d_out may equal d_m.  The result seeds icon_reconstruct's
 * d_m_init. */
int icon_unet_warm_start(icon_unet* net, int S, const float* d_m, float* d_out, void* stream);
/* Factor g on the warm start's result: d_out = g * (d_m + s * U-Net(d_m / s)).
 * Default 1 (the result exactly as above).  A learned-start solver chooses it
 * with the continuation length on held-out data: the descent iterations that
 * follow reach a given data residual sooner from a start a little below the
 * network's amplitude. */
int icon_unet_set_output_gain(icon_unet* net, float gain);
/* Gradients of loss = mean (U-Net(in) - target)^2 with respect to the
 * parameters (training net); *loss receives the loss. */
int icon_unet_gradients(icon_unet* net, int B, const float* d_in, const float* d_target, double* loss, void* stream);
int icon_unet_get_gradients(const icon_unet* net, float* host, int64_t count);
/* One Adam step (beta 0.9 / 0.999) with learning rate lr. */
int icon_unet_train_step(icon_unet* net, int B, const float* d_in, const float* d_target, float lr, double* loss,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ICON_UNET_H */
