// Learned warm start: a 4-level 2D U-Net (32/64/128/256 channels) trained on
// pairs of early-iteration and converged ICON slices, run on BF16 MFMA
// kernels with FP32 accumulation (unet_kernels.hip).  No counterpart in the
// reference; its output seeds icon_reconstruct's m_init.
//
// Architecture (NHWC, ReLU after every 3x3 convolution):
//   e1 = conv(1->32), conv(32->32)            H
//   e2 = pool, conv(32->64), conv(64->64)     H/2
//   e3 = pool, conv(64->128), conv(128->128)  H/4
//   b  = pool, conv(128->256), conv(256->256) H/8
//   d3 = up(b) ++ e3 -> conv(384->128), conv(128->128)
//   d2 = up(d3) ++ e2 -> conv(192->64), conv(64->64)
//   d1 = up(d2) ++ e1 -> conv(96->32), conv(32->32)
//   out = 1x1 conv(32->1)
// pool: 2x2 average; up: nearest x2; "++": channel concatenation, realised by
// writing the encoder output and the upsampled decoder input into one buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icon_unet.h"
#include "unet_kernels.h"

using namespace unet;

namespace {

thread_local std::string g_uerr;

int ufail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_uerr = buf;
  return code;
}

#define UCHK(expr)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return ufail(2, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

struct Conv {
  int cin, cout, taps, cin_true;
  long off_w, off_b;       // offsets into the flat parameter vector
  uint16_t* w16 = nullptr;   // [taps][cout][cin]
  uint16_t* wt16 = nullptr;  // flipped, for input gradients
};

// bf16 buffer
struct T16 {
  uint16_t* p = nullptr;
  int C = 0;
};
struct T32 {
  float* p = nullptr;
  int C = 0;
};

}  // namespace

struct icon_unet {
  int device = 0;
  int maxB = 0, H = 0, W = 0;
  bool train = false;
  std::vector<Conv> convs;   // 14 3x3 convolutions
  long head_w = 0, head_b = 0;
  long nparam = 0;
  float* params = nullptr;   // fp32 master
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long adam_t = 0;
  double* d_loss = nullptr;
  float* gpart = nullptr;    // training: partial sums of the weight, bias and head gradients
  // activations (bf16)
  T16 x8, a1, cat1, p1, a2, cat2, p2, a3, cat3, p3, ab, bb, d3a, d3b, d2a, d2b, d1a, d1b;
  float* out = nullptr;      // [maxB][H][W] network output of the warm start
  float* scale = nullptr;    // [maxB] per-image input scale of the warm start
  float out_gain = 1.f;      // factor on the warm start's output (icon_unet_set_output_gain)
  double* rms_part = nullptr;  // [maxB][image_rms_parts()] partial sums of the scale
  // gradients (fp32)
  T32 g_d1b, g_d1a, g_cat1, g_d2b, g_d2a, g_cat2, g_d3b, g_d3a, g_cat3, g_bb, g_ab, g_p3, g_a3, g_p2, g_a2, g_p1,
      g_a1, g_out;
  std::vector<void*> allocs;
};

namespace {

enum { E1A, E1B, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1A, D1B, NCONV };

int alloc_bytes(icon_unet* n, void** p, size_t bytes) {
  UCHK(hipMalloc(p, bytes ? bytes : 16));
  UCHK(hipMemset(*p, 0, bytes ? bytes : 16));
  n->allocs.push_back(*p);
  return 0;
}

int t16(icon_unet* n, T16& t, long px, int C) {
  t.C = C;
  return alloc_bytes(n, (void**)&t.p, sizeof(uint16_t) * px * C);
}
int t32(icon_unet* n, T32& t, long px, int C) {
  t.C = C;
  return alloc_bytes(n, (void**)&t.p, sizeof(float) * px * C);
}

// He-normal initial weights from a seeded xorshift / Box-Muller stream.
struct Rng {
  uint64_t s;
  double next() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return ((s >> 11) + 0.5) / 9007199254740992.0;
  }
  double normal() {
    const double u = next(), v = next();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
  }
};

int refresh_bf16(icon_unet* n, hipStream_t st) {
  // bf16 copies of the master weights (through the Adam kernel with zero lr
  // would also touch m, v; use the plain conversion kernel instead)
  for (auto& c : n->convs) {
    const long nw = (long)c.taps * c.cout * c.cin;
    UCHK(launch_f32_to_bf16(n->params + c.off_w, c.w16, nw, 1, 0, 1, st));
    UCHK(launch_flip_weights(c.w16, c.wt16, c.taps, c.cout, c.cin, st));
  }
  return 0;
}

// The network's first convolution computed inside the second one's input
// staging (from the FP32 images, divided by scale when given); its output
// written to a1_out when given.
struct FirstConv {
  const float* img;
  const float* scale;
  uint16_t* a1_out;
};

// up (optional): the input's first Cup channels, read from this half-size
// tensor (nearest-neighbour 2x upsampling) instead of x; pool (optional):
// the output's 2x2 average, written at half size
template <typename TX, typename TY>
int conv(icon_unet* n, const Conv& c, const TX* x, int ldx, int cx0, TY* y, int ldy, int cy0, int B, int H, int W,
         bool relu, const uint16_t* mask, int ldm, int cm0, bool flipped, bool accumulate, hipStream_t st,
         const T16* up = nullptr, const T16* pool = nullptr, const float* head_w = nullptr,
         const float* head_b = nullptr, float* head_out = nullptr, const FirstConv* first = nullptr) {
  ConvArgs<TX, TY> a{};
  a.x = x; a.ldx = ldx; a.cx0 = cx0;
  a.w = flipped ? c.wt16 : c.w16;
  a.bias = flipped ? nullptr : n->params + c.off_b;
  a.y = y; a.ldy = ldy; a.cy0 = cy0;
  a.mask = mask; a.ldm = ldm; a.cm0 = cm0;
  a.N = B; a.H = H; a.W = W;
  a.Cin = flipped ? c.cout : c.cin;
  a.Cout = flipped ? c.cin : c.cout;
  a.taps = c.taps;
  a.relu = relu; a.accumulate = accumulate;
  if (up) { a.up = up->p; a.ldu = up->C; a.cu0 = 0; a.Cup = up->C; }
  if (pool) { a.pool = pool->p; a.ldp = pool->C; a.cp0 = 0; }
  a.head_w = head_w; a.head_b = head_b; a.head_out = head_out;
  if constexpr (sizeof(TX) == 2 && sizeof(TY) == 2) {
    if (first) {
      const Conv& c1 = n->convs[0];
      a.x1 = first->img; a.x1_scale = first->scale; a.a1_out = first->a1_out;
      a.w1 = c1.w16; a.w1_cin = c1.cin; a.b1 = n->params + c1.off_b;
    }
  }
  UCHK(launch_conv(a, st));
  return 0;
}

// pixel splits of a weight gradient over P pixels
int wgrad_splits(long P) { return (int)std::max<long>(1, std::min<long>(256, P / 2048)); }

int wgrad(icon_unet* n, const Conv& c, const uint16_t* x, int ldx, int cx0, const float* dy, int lddy, int cdy0,
          int B, int H, int W, hipStream_t st) {
  WgradArgs a{};
  a.x = x; a.ldx = ldx; a.cx0 = cx0; a.dy = dy; a.lddy = lddy; a.cdy0 = cdy0;
  a.dw = n->grads + c.off_w;
  a.N = B; a.H = H; a.W = W; a.Cin = c.cin; a.Cout = c.cout; a.taps = c.taps;
  a.part = n->gpart;
  const long P = (long)B * H * W;
  UCHK(launch_wgrad(a, wgrad_splits(P), st));
  UCHK(launch_bias_grad(dy, lddy, cdy0, P, c.cout, n->grads + c.off_b, n->gpart, st));
  return 0;
}

int up(const T16& in, const T16& out, int co0, int B, int h, int w, hipStream_t st) {
  PoolArgs a{in.p, in.C, 0, out.p, out.C, co0, B, h, w, in.C};
  UCHK(launch_up_fwd(a, st));
  return 0;
}

// in_scale (optional): per-image divisor of the input
int forward(icon_unet* n, int B, const float* d_in, float* d_out, hipStream_t st, const float* in_scale = nullptr) {
  const int H = n->H, W = n->W, H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4, H8 = H / 8, W8 = W / 8;
  const long P = (long)B * H * W;
  auto& cv = n->convs;
  int rc;
#define R(x) do { if ((rc = (x))) return rc; } while (0)
  if (n->train) {
    // the BF16 input for the first convolution's weight gradient
    const hipError_t ce = in_scale ? launch_to_bf16_scaled(d_in, in_scale, (long)H * W, B, n->x8.p, st)
                                   : launch_f32_to_bf16(d_in, n->x8.p, P, 8, 0, 1, st);
    if (ce != hipSuccess) return ufail(2, "input conversion: %s", hipGetErrorString(ce));
  }
  // the first convolution (E1A) inside E1B's input staging; its output is
  // stored for the backward pass only
  const FirstConv first{d_in, in_scale, n->train ? n->a1.p : nullptr};
  // the encoder convolutions also write their 2x2 average pool; without
  // training (no weight gradients need the concatenated decoder inputs) the
  // decoder convolutions read the upsampled half of their input straight
  // from the half-size tensor
  const bool fuse_up = !n->train;
  // the skip halves: channels [2C, 3C) of cat in training, all of cat
  // otherwise; a decoder reads its input channel ci >= Cup (the skip half)
  // at channel ci - u of cat, u = Cup when only the skip half is stored
  const int k1 = fuse_up ? 0 : 64, k2 = fuse_up ? 0 : 128, k3 = fuse_up ? 0 : 256;
  const int u1 = fuse_up ? 64 : 0, u2 = fuse_up ? 128 : 0, u3 = fuse_up ? 256 : 0;
  R(conv(n, cv[E1B], n->a1.p, 32, 0, n->cat1.p, n->cat1.C, k1, B, H, W, true, nullptr, 0, 0, false, false, st, nullptr,
         &n->p1, nullptr, nullptr, nullptr, &first));
  R(conv(n, cv[E2A], n->p1.p, 32, 0, n->a2.p, 64, 0, B, H2, W2, true, nullptr, 0, 0, false, false, st));
  R(conv(n, cv[E2B], n->a2.p, 64, 0, n->cat2.p, n->cat2.C, k2, B, H2, W2, true, nullptr, 0, 0, false, false, st, nullptr,
         &n->p2));
  R(conv(n, cv[E3A], n->p2.p, 64, 0, n->a3.p, 128, 0, B, H4, W4, true, nullptr, 0, 0, false, false, st));
  R(conv(n, cv[E3B], n->a3.p, 128, 0, n->cat3.p, n->cat3.C, k3, B, H4, W4, true, nullptr, 0, 0, false, false, st, nullptr,
         &n->p3));
  R(conv(n, cv[BA], n->p3.p, 128, 0, n->ab.p, 256, 0, B, H8, W8, true, nullptr, 0, 0, false, false, st));
  R(conv(n, cv[BB], n->ab.p, 256, 0, n->bb.p, 256, 0, B, H8, W8, true, nullptr, 0, 0, false, false, st));
  if (!fuse_up) R(up(n->bb, n->cat3, 0, B, H8, W8, st));
  R(conv(n, cv[D3A], n->cat3.p, n->cat3.C, -u3, n->d3a.p, 128, 0, B, H4, W4, true, nullptr, 0, 0, false, false, st,
         fuse_up ? &n->bb : nullptr));
  R(conv(n, cv[D3B], n->d3a.p, 128, 0, n->d3b.p, 128, 0, B, H4, W4, true, nullptr, 0, 0, false, false, st));
  if (!fuse_up) R(up(n->d3b, n->cat2, 0, B, H4, W4, st));
  R(conv(n, cv[D2A], n->cat2.p, n->cat2.C, -u2, n->d2a.p, 64, 0, B, H2, W2, true, nullptr, 0, 0, false, false, st,
         fuse_up ? &n->d3b : nullptr));
  R(conv(n, cv[D2B], n->d2a.p, 64, 0, n->d2b.p, 64, 0, B, H2, W2, true, nullptr, 0, 0, false, false, st));
  if (!fuse_up) R(up(n->d2b, n->cat1, 0, B, H2, W2, st));
  R(conv(n, cv[D1A], n->cat1.p, n->cat1.C, -u1, n->d1a.p, 32, 0, B, H, W, true, nullptr, 0, 0, false, false, st,
         fuse_up ? &n->d2b : nullptr));
  if (n->train) {
    R(conv(n, cv[D1B], n->d1a.p, 32, 0, n->d1b.p, 32, 0, B, H, W, true, nullptr, 0, 0, false, false, st));
    if (launch_head_fwd(n->d1b.p, 32, n->params + n->head_w, n->params + n->head_b, 32, P, d_out, st) != hipSuccess)
      return ufail(2, "head");
  } else {
    // inference: the head in D1B's epilogue (the same sums on the same BF16
    // values), D1B's output is not stored
    R(conv(n, cv[D1B], n->d1a.p, 32, 0, n->d1b.p, 32, 0, B, H, W, true, nullptr, 0, 0, false, false, st, nullptr,
           nullptr, n->params + n->head_w, n->params + n->head_b, d_out));
  }
#undef R
  return 0;
}

int backward(icon_unet* n, int B, const float* d_dout, hipStream_t st) {
  const int H = n->H, W = n->W, H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4, H8 = H / 8, W8 = W / 8;
  const long P = (long)B * H * W;
  auto& cv = n->convs;
  int rc;
#define R(x) do { if ((rc = (x))) return rc; } while (0)
#define HIPR(x) do { if ((x) != hipSuccess) return ufail(2, "%s failed", #x); } while (0)
  HIPR(launch_head_bwd(n->d1b.p, 32, n->params + n->head_w, 32, P, d_dout, n->g_d1b.p, 32, n->grads + n->head_w,
                       n->grads + n->head_b, n->gpart, st));
  // d1b, d1a
  R(wgrad(n, cv[D1B], n->d1a.p, 32, 0, n->g_d1b.p, 32, 0, B, H, W, st));
  R(conv(n, cv[D1B], n->g_d1b.p, 32, 0, n->g_d1a.p, 32, 0, B, H, W, false, n->d1a.p, 32, 0, true, false, st));
  R(wgrad(n, cv[D1A], n->cat1.p, 96, 0, n->g_d1a.p, 32, 0, B, H, W, st));
  R(conv(n, cv[D1A], n->g_d1a.p, 32, 0, n->g_cat1.p, 96, 0, B, H, W, false, n->cat1.p, 96, 0, true, false, st));
  {
    PoolGradArgs u{n->g_cat1.p, 96, 0, n->g_d2b.p, 64, 0, B, H2, W2, 64, 0};
    HIPR(launch_up_bwd(u, st));
  }
  // d2b, d2a
  R(wgrad(n, cv[D2B], n->d2a.p, 64, 0, n->g_d2b.p, 64, 0, B, H2, W2, st));
  R(conv(n, cv[D2B], n->g_d2b.p, 64, 0, n->g_d2a.p, 64, 0, B, H2, W2, false, n->d2a.p, 64, 0, true, false, st));
  R(wgrad(n, cv[D2A], n->cat2.p, 192, 0, n->g_d2a.p, 64, 0, B, H2, W2, st));
  R(conv(n, cv[D2A], n->g_d2a.p, 64, 0, n->g_cat2.p, 192, 0, B, H2, W2, false, n->cat2.p, 192, 0, true, false, st));
  {
    PoolGradArgs u{n->g_cat2.p, 192, 0, n->g_d3b.p, 128, 0, B, H4, W4, 128, 0};
    HIPR(launch_up_bwd(u, st));
  }
  // d3b, d3a
  R(wgrad(n, cv[D3B], n->d3a.p, 128, 0, n->g_d3b.p, 128, 0, B, H4, W4, st));
  R(conv(n, cv[D3B], n->g_d3b.p, 128, 0, n->g_d3a.p, 128, 0, B, H4, W4, false, n->d3a.p, 128, 0, true, false, st));
  R(wgrad(n, cv[D3A], n->cat3.p, 384, 0, n->g_d3a.p, 128, 0, B, H4, W4, st));
  R(conv(n, cv[D3A], n->g_d3a.p, 128, 0, n->g_cat3.p, 384, 0, B, H4, W4, false, n->cat3.p, 384, 0, true, false, st));
  {
    PoolGradArgs u{n->g_cat3.p, 384, 0, n->g_bb.p, 256, 0, B, H8, W8, 256, 0};
    HIPR(launch_up_bwd(u, st));
  }
  // bottleneck
  R(wgrad(n, cv[BB], n->ab.p, 256, 0, n->g_bb.p, 256, 0, B, H8, W8, st));
  R(conv(n, cv[BB], n->g_bb.p, 256, 0, n->g_ab.p, 256, 0, B, H8, W8, false, n->ab.p, 256, 0, true, false, st));
  R(wgrad(n, cv[BA], n->p3.p, 128, 0, n->g_ab.p, 256, 0, B, H8, W8, st));
  R(conv(n, cv[BA], n->g_ab.p, 256, 0, n->g_p3.p, 128, 0, B, H8, W8, false, nullptr, 0, 0, true, false, st));
  {
    PoolGradArgs u{n->g_p3.p, 128, 0, n->g_cat3.p, 384, 256, B, H4, W4, 128, 1};
    HIPR(launch_pool_bwd(u, st));
    HIPR(launch_relu_mask(n->g_cat3.p, 384, 256, n->cat3.p, 384, 256, (long)B * H4 * W4, 128, st));
  }
  // e3
  R(wgrad(n, cv[E3B], n->a3.p, 128, 0, n->g_cat3.p, 384, 256, B, H4, W4, st));
  R(conv(n, cv[E3B], n->g_cat3.p, 384, 256, n->g_a3.p, 128, 0, B, H4, W4, false, n->a3.p, 128, 0, true, false, st));
  R(wgrad(n, cv[E3A], n->p2.p, 64, 0, n->g_a3.p, 128, 0, B, H4, W4, st));
  R(conv(n, cv[E3A], n->g_a3.p, 128, 0, n->g_p2.p, 64, 0, B, H4, W4, false, nullptr, 0, 0, true, false, st));
  {
    PoolGradArgs u{n->g_p2.p, 64, 0, n->g_cat2.p, 192, 128, B, H2, W2, 64, 1};
    HIPR(launch_pool_bwd(u, st));
    HIPR(launch_relu_mask(n->g_cat2.p, 192, 128, n->cat2.p, 192, 128, (long)B * H2 * W2, 64, st));
  }
  // e2
  R(wgrad(n, cv[E2B], n->a2.p, 64, 0, n->g_cat2.p, 192, 128, B, H2, W2, st));
  R(conv(n, cv[E2B], n->g_cat2.p, 192, 128, n->g_a2.p, 64, 0, B, H2, W2, false, n->a2.p, 64, 0, true, false, st));
  R(wgrad(n, cv[E2A], n->p1.p, 32, 0, n->g_a2.p, 64, 0, B, H2, W2, st));
  R(conv(n, cv[E2A], n->g_a2.p, 64, 0, n->g_p1.p, 32, 0, B, H2, W2, false, nullptr, 0, 0, true, false, st));
  {
    PoolGradArgs u{n->g_p1.p, 32, 0, n->g_cat1.p, 96, 64, B, H, W, 32, 1};
    HIPR(launch_pool_bwd(u, st));
    HIPR(launch_relu_mask(n->g_cat1.p, 96, 64, n->cat1.p, 96, 64, P, 32, st));
  }
  // e1
  R(wgrad(n, cv[E1B], n->a1.p, 32, 0, n->g_cat1.p, 96, 64, B, H, W, st));
  R(conv(n, cv[E1B], n->g_cat1.p, 96, 64, n->g_a1.p, 32, 0, B, H, W, false, n->a1.p, 32, 0, true, false, st));
  R(wgrad(n, cv[E1A], n->x8.p, 8, 0, n->g_a1.p, 32, 0, B, H, W, st));
#undef R
#undef HIPR
  return 0;
}

}  // namespace

extern "C" {

const char* icon_unet_last_error(void) { return g_uerr.c_str(); }

int icon_unet_create(icon_unet** out, int device, int max_batch, int H, int W, int train) {
  if (!out || max_batch < 1 || H < 8 || W < 8 || H % 8 || W % 8)
    return ufail(1, "icon_unet_create: need max_batch >= 1 and H, W multiples of 8");
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return ufail(1, "icon_unet_create: no device %d", device);
  auto* n = new icon_unet();
  n->device = device; n->maxB = max_batch; n->H = H; n->W = W; n->train = train != 0;
  const int spec[NCONV][3] = {{8, 32, 1},    {32, 32, 32},   {32, 64, 32},   {64, 64, 64},   {64, 128, 64},
                              {128, 128, 128}, {128, 256, 128}, {256, 256, 256}, {384, 128, 384}, {128, 128, 128},
                              {192, 64, 192}, {64, 64, 64},    {96, 32, 96},    {32, 32, 32}};
  long off = 0;
  for (int i = 0; i < NCONV; i++) {
    Conv c;
    c.cin = spec[i][0]; c.cout = spec[i][1]; c.cin_true = spec[i][2]; c.taps = 9;
    c.off_w = off; off += (long)c.taps * c.cout * c.cin;
    c.off_b = off; off += c.cout;
    n->convs.push_back(c);
  }
  n->head_w = off; off += 32;
  n->head_b = off; off += 1;
  n->nparam = off;
  int rc = 0;
  auto bail = [&](int code) { icon_unet_destroy(n); return code; };
  if ((rc = alloc_bytes(n, (void**)&n->params, sizeof(float) * off))) return bail(rc);
  for (auto& c : n->convs) {
    const long nw = (long)c.taps * c.cout * c.cin;
    if ((rc = alloc_bytes(n, (void**)&c.w16, sizeof(uint16_t) * nw))) return bail(rc);
    if ((rc = alloc_bytes(n, (void**)&c.wt16, sizeof(uint16_t) * nw))) return bail(rc);
  }
  if ((rc = alloc_bytes(n, (void**)&n->d_loss, sizeof(double)))) return bail(rc);
  const long P = (long)max_batch * H * W, P2 = P / 4, P4 = P / 16, P8 = P / 64;
  if ((rc = alloc_bytes(n, (void**)&n->out, sizeof(float) * P)) ||
      (rc = alloc_bytes(n, (void**)&n->scale, sizeof(float) * max_batch)) ||
      (rc = alloc_bytes(n, (void**)&n->rms_part, sizeof(double) * max_batch * image_rms_parts())))
    return bail(rc);
  // cat1..3: the decoders' concatenated inputs [upsampled | skip] for
  // training; without training the upsampled half is read from the
  // half-size tensor and only the skip half is stored, densely
  const int sk = n->train ? 3 : 1;
  // x8 (the BF16 input) and a1 (the first convolution) only for training
  const long Pt = n->train ? P : 1;
  if ((rc = t16(n, n->x8, Pt, 8)) || (rc = t16(n, n->a1, Pt, 32)) || (rc = t16(n, n->cat1, P, 32 * sk)) ||
      (rc = t16(n, n->p1, P2, 32)) || (rc = t16(n, n->a2, P2, 64)) || (rc = t16(n, n->cat2, P2, 64 * sk)) ||
      (rc = t16(n, n->p2, P4, 64)) || (rc = t16(n, n->a3, P4, 128)) || (rc = t16(n, n->cat3, P4, 128 * sk)) ||
      (rc = t16(n, n->p3, P8, 128)) || (rc = t16(n, n->ab, P8, 256)) || (rc = t16(n, n->bb, P8, 256)) ||
      (rc = t16(n, n->d3a, P4, 128)) || (rc = t16(n, n->d3b, P4, 128)) || (rc = t16(n, n->d2a, P2, 64)) ||
      (rc = t16(n, n->d2b, P2, 64)) || (rc = t16(n, n->d1a, P, 32)) || (rc = t16(n, n->d1b, P, 32)))
    return bail(rc);
  if (n->train) {
    // gradient partial sums: the largest of a weight gradient's splits x
    // weights (at the level of each convolution), the bias and head blocks
    long part = (long)kHeadGradParts * 33;
    const long level_px[NCONV] = {P, P, P2, P2, P4, P4, P8, P8, P4, P4, P2, P2, P, P};
    for (int i = 0; i < NCONV; i++) {
      const Conv& c = n->convs[i];
      part = std::max(part, (long)wgrad_splits(level_px[i]) * c.taps * c.cout * c.cin);
      part = std::max(part, (long)kBiasGradParts * c.cout);
    }
    if ((rc = alloc_bytes(n, (void**)&n->gpart, sizeof(float) * part))) return bail(rc);
    if ((rc = alloc_bytes(n, (void**)&n->grads, sizeof(float) * off)) ||
        (rc = alloc_bytes(n, (void**)&n->adam_m, sizeof(float) * off)) ||
        (rc = alloc_bytes(n, (void**)&n->adam_v, sizeof(float) * off)))
      return bail(rc);
    if ((rc = t32(n, n->g_d1b, P, 32)) || (rc = t32(n, n->g_d1a, P, 32)) || (rc = t32(n, n->g_cat1, P, 96)) ||
        (rc = t32(n, n->g_d2b, P2, 64)) || (rc = t32(n, n->g_d2a, P2, 64)) || (rc = t32(n, n->g_cat2, P2, 192)) ||
        (rc = t32(n, n->g_d3b, P4, 128)) || (rc = t32(n, n->g_d3a, P4, 128)) || (rc = t32(n, n->g_cat3, P4, 384)) ||
        (rc = t32(n, n->g_bb, P8, 256)) || (rc = t32(n, n->g_ab, P8, 256)) || (rc = t32(n, n->g_p3, P8, 128)) ||
        (rc = t32(n, n->g_a3, P4, 128)) || (rc = t32(n, n->g_p2, P4, 64)) || (rc = t32(n, n->g_a2, P2, 64)) ||
        (rc = t32(n, n->g_p1, P2, 32)) || (rc = t32(n, n->g_a1, P, 32)) || (rc = t32(n, n->g_out, P, 1)))
      return bail(rc);
  }
  *out = n;
  return 0;
}

int icon_unet_destroy(icon_unet* n) {
  if (!n) return 0;
  (void)hipDeviceSynchronize();
  for (void* p : n->allocs) (void)hipFree(p);
  delete n;
  return 0;
}

int icon_unet_param_count(const icon_unet* n, int64_t* count) {
  if (!n || !count) return ufail(1, "null");
  *count = n->nparam;
  return 0;
}

int icon_unet_init(icon_unet* n, uint64_t seed) {
  if (!n) return ufail(1, "null");
  std::vector<float> h(n->nparam, 0.f);
  Rng rng{seed * 2654435761ull + 88172645463325252ull};
  for (auto& c : n->convs) {
    const double sd = std::sqrt(2.0 / (c.taps * c.cin_true));
    for (int t = 0; t < c.taps; t++)
      for (int co = 0; co < c.cout; co++)
        for (int ci = 0; ci < c.cin; ci++)
          h[c.off_w + ((long)t * c.cout + co) * c.cin + ci] = ci < c.cin_true ? (float)(sd * rng.normal()) : 0.f;
  }
  for (int ci = 0; ci < 32; ci++) h[n->head_w + ci] = (float)(std::sqrt(1.0 / 32) * rng.normal());
  return icon_unet_set_params(n, h.data(), n->nparam);
}

int icon_unet_get_params(const icon_unet* n, float* host, int64_t count) {
  if (!n || !host || count != n->nparam) return ufail(1, "icon_unet_get_params: count must be %ld", n ? n->nparam : 0L);
  UCHK(hipMemcpy(host, n->params, sizeof(float) * count, hipMemcpyDeviceToHost));
  return 0;
}

int icon_unet_set_params(icon_unet* n, const float* host, int64_t count) {
  if (!n || !host || count != n->nparam) return ufail(1, "icon_unet_set_params: count must be %ld", n ? n->nparam : 0L);
  UCHK(hipMemcpy(n->params, host, sizeof(float) * count, hipMemcpyHostToDevice));
  int rc = refresh_bf16(n, nullptr);
  if (rc) return rc;
  UCHK(hipDeviceSynchronize());
  n->adam_t = 0;
  if (n->train) {
    UCHK(hipMemset(n->adam_m, 0, sizeof(float) * count));
    UCHK(hipMemset(n->adam_v, 0, sizeof(float) * count));
  }
  return 0;
}

int icon_unet_forward(icon_unet* n, int B, const float* d_in, float* d_out, void* stream) {
  if (!n || !d_in || !d_out || B < 1 || B > n->maxB) return ufail(1, "icon_unet_forward: bad argument");
  return forward(n, B, d_in, d_out, (hipStream_t)stream);
}

int icon_unet_warm_start(icon_unet* n, int S, const float* d_m, float* d_out, void* stream) {
  if (!n || !d_m || !d_out || S < 1) return ufail(1, "icon_unet_warm_start: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)n->H * n->W;
  for (int s0 = 0; s0 < S; s0 += n->maxB) {
    const int B = std::min(n->maxB, S - s0);
    const float* m = d_m + (long)s0 * P;
    UCHK(launch_image_rms(m, P, B, n->rms_part, n->scale, st));
    int rc = forward(n, B, m, n->out, st, n->scale);
    if (rc) return rc;
    UCHK(launch_residual_add(m, n->out, n->scale, P, B, n->out_gain, d_out + (long)s0 * P, st));
  }
  return 0;
}

int icon_unet_set_output_gain(icon_unet* n, float gain) {
  if (!n || !(gain > 0.f) || !(gain < 16.f)) return ufail(1, "icon_unet_set_output_gain: need 0 < gain < 16");
  n->out_gain = gain;
  return 0;
}

int icon_unet_gradients(icon_unet* n, int B, const float* d_in, const float* d_target, double* loss, void* stream) {
  if (!n || !n->train || !d_in || !d_target || B < 1 || B > n->maxB)
    return ufail(1, "icon_unet_gradients: bad argument (training net required)");
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)B * n->H * n->W;
  int rc = forward(n, B, d_in, n->g_out.p, st);   // prediction into g_out, then its gradient in place
  if (rc) return rc;
  UCHK(launch_zero32(n->grads, n->nparam, st));
  UCHK(hipMemsetAsync(n->d_loss, 0, sizeof(double), st));
  // loss = sum (pred - target)^2 / P; dL/dpred = 2 (pred - target) / P
  UCHK(launch_mse(n->g_out.p, d_target, P, n->g_out.p, n->d_loss, 2.f / (float)P, st));
  if ((rc = backward(n, B, n->g_out.p, st))) return rc;
  if (loss) {
    double l = 0;
    UCHK(hipMemcpyAsync(&l, n->d_loss, sizeof(double), hipMemcpyDeviceToHost, st));
    UCHK(hipStreamSynchronize(st));
    *loss = l / (double)P;
  }
  return 0;
}

int icon_unet_get_gradients(const icon_unet* n, float* host, int64_t count) {
  if (!n || !n->train || !host || count != n->nparam) return ufail(1, "icon_unet_get_gradients: bad argument");
  UCHK(hipMemcpy(host, n->grads, sizeof(float) * count, hipMemcpyDeviceToHost));
  return 0;
}

int icon_unet_train_step(icon_unet* n, int B, const float* d_in, const float* d_target, float lr, double* loss,
                         void* stream) {
  int rc = icon_unet_gradients(n, B, d_in, d_target, loss, stream);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  n->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)n->adam_t)), c2 = 1.f / (1.f - std::pow(b2, (float)n->adam_t));
  UCHK(launch_adam(n->params, nullptr, n->grads, n->adam_m, n->adam_v, n->nparam, lr, b1, b2, 1e-8f, c1, c2, 1.f, st));
  return refresh_bf16(n, st);
}

}  // extern "C"
