// Argument blocks and launchers of the U-Net kernels (NHWC tensors; bf16 as uint16).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace unet {

// Convolution y = conv(x, w) (+ bias, ReLU), taps 9 (3x3, zero padding 1) or 1.
// x: [N][H][W] pixels, channel stride ldx, channels [cx0, cx0 + Cin);
// y: channel stride ldy, channels [cy0, cy0 + Cout); w: [taps][Cout][Cin] bf16.
// mask (optional): y is zeroed where the bf16 mask tensor is <= 0 (the ReLU of
// a forward activation, for input gradients); accumulate: y += result.
template <typename TX, typename TY>
struct ConvArgs {
  const TX* x;
  int ldx, cx0;
  const uint16_t* w;
  const float* bias;
  TY* y;
  int ldy, cy0;
  const uint16_t* mask;
  int ldm, cm0;
  int N, H, W, Cin, Cout, taps;
  int relu, accumulate;
  // optional: input channels [0, Cup) read from `up` (H/2 x W/2 pixels,
  // nearest-neighbour 2x upsampling) instead of x
  const uint16_t* up;
  int ldu, cu0, Cup;
  // optional (BF16 output): 2x2 average pool of the output, written to pool
  // (H/2 x W/2 pixels)
  uint16_t* pool;
  int ldp, cp0;
  // optional (BF16 output, Cout = 32, inference): the 1x1 head fused into the
  // epilogue, head_out[p] = head_b + sum_c head_w[c] y[p][c] over the BF16
  // values of y in channel order (what k_head_fwd computes); y itself is not
  // stored
  const float* head_w;
  const float* head_b;
  float* head_out;
  // optional (BF16 input and output, Cin = Cout = 32, 3x3, no up): x is not
  // read but computed per tile as the network's first convolution of the
  // one-channel FP32 images x1: x[p][c] = relu(b1[c] + sum_t w1[t][c][0]
  // img[p + off(t)]), img = bf16(x1 / x1_scale[image]) (x1_scale optional)
  // and zero outside the image, w1 [9][32][w1_cin] BF16; written to a1_out
  // (32 channels per pixel) when given
  const float* x1;
  const float* x1_scale;
  const uint16_t* w1;
  int w1_cin;
  const float* b1;
  uint16_t* a1_out;
};

// dw[tap][co][ci] += sum over pixels of dy[p][co] * x[p + offset(tap)][ci]
struct WgradArgs {
  const uint16_t* x;
  int ldx, cx0;
  const float* dy;
  int lddy, cdy0;
  float* dw;
  int N, H, W, Cin, Cout, taps;
  float* part;   // [splits][taps][Cout][Cin] scratch: each pixel split's sums, added to dw in split order
};

// 2x2 average pool (fwd: in H x W -> out H/2 x W/2) or nearest 2x upsample
// (fwd: in H x W -> out 2H x 2W); bf16 with channel sub-ranges.
struct PoolArgs {
  const uint16_t* in;
  int ldi, ci0;
  uint16_t* out;
  int ldo, co0;
  int N, H, W, C;   // input dims
};

// Gradients of pool / upsample (fp32).  H, W are the dims of din.
struct PoolGradArgs {
  const float* dout;
  int ldo, co0;
  float* din;
  int ldi, ci0;
  int N, H, W, C;
  int accumulate;
};

template <typename TX, typename TY>
hipError_t launch_conv(const ConvArgs<TX, TY>& a, hipStream_t st);
hipError_t launch_wgrad(const WgradArgs& a, int splits, hipStream_t st);
// Weight and bias gradients are reduced without atomics: every block writes
// its partial sums to a scratch buffer and launch_sum_parts adds them to the
// gradient in a fixed order, so that training is reproducible run to run.
hipError_t launch_bias_grad(const float* dy, int ld, int c0, long P, int C, float* db, float* part, hipStream_t st);
constexpr int kBiasGradParts = 256;   // partial sums per channel (launch_bias_grad's part: [256][C])
constexpr int kHeadGradParts = 256;   // blocks of launch_head_bwd (its part: [256][C + 1])
// out[i] += part[0][i] + part[1][i] + ... + part[nparts - 1][i] (in that order)
hipError_t launch_sum_parts(const float* part, int nparts, long n, float* out, hipStream_t st);
hipError_t launch_pool_bwd(const PoolGradArgs& a, hipStream_t st);
hipError_t launch_up_fwd(const PoolArgs& a, hipStream_t st);
hipError_t launch_up_bwd(const PoolGradArgs& a, hipStream_t st);
hipError_t launch_relu_mask(float* g, int ldg, int cg0, const uint16_t* act, int lda, int ca0, long P, int C,
                            hipStream_t st);
hipError_t launch_head_fwd(const uint16_t* x, int ldx, const float* w, const float* b, int C, long P, float* out,
                           hipStream_t st);
hipError_t launch_head_bwd(const uint16_t* x, int ldx, const float* w, int C, long P, const float* dout, float* dx,
                           int lddx, float* dw, float* db, float* part, hipStream_t st);
hipError_t launch_flip_weights(const uint16_t* w, uint16_t* wt, int taps, int Cout, int Cin, hipStream_t st);
hipError_t launch_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale, hipStream_t st);
hipError_t launch_f32_to_bf16(const float* in, uint16_t* out, long n, int ldo, int co0, int C, hipStream_t st);
hipError_t launch_mse(const float* pred, const float* target, long P, float* dpred, double* loss, float scale,
                      hipStream_t st);
// scale[b] = RMS of image b (1 if zero); part holds B * image_rms_parts() doubles
hipError_t launch_image_rms(const float* in, long P, int B, double* part, float* scale, hipStream_t st);
int image_rms_parts();
hipError_t launch_to_bf16_scaled(const float* in, const float* scale, long P, int B, uint16_t* out, hipStream_t st);
// out = gain * (m + scale[b] * y) per image b (gain 1: m + scale[b] * y exactly)
hipError_t launch_residual_add(const float* m, const float* y, const float* scale, long P, int B, float gain,
                               float* out, hipStream_t st);
hipError_t launch_zero32(float* p, long n, hipStream_t st);

}  // namespace unet
