// CDNA4 kernels of the learned warm start: a 2D U-Net on BF16 MFMA.
//
// Tensors are NHWC.  Forward activations are BF16 (stored as uint16), the
// gradients FP32; weights keep an FP32 master copy and a BF16 compute copy in
// the layout [tap][c_out][c_in] (c_in fastest), taps ky*3 + kx.
//
// The 3x3 (and 1x1) convolution is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16: M = output channels, N = pixels, K = taps x
// input channels in groups of 32.  Lane l feeds A[channel slot l%16][8 input
// channels 8*(l/16)...] (weights) and B[8 input channels 8*(l/16)...][pixel
// l%16] as 16-byte vectors from LDS-staged tiles and receives D[slots
// 4*(l/16)+i][pixel l%16], the slots ordered so that they are consecutive
// output channels.  The same kernel computes the input gradient with the
// spatially flipped, channel-transposed weights; the weight gradient is a
// split-K GEMM over pixels whose operands are transposed through LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "unet_kernels.h"

namespace unet {

typedef __attribute__((ext_vector_type(4))) float f4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf8;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);   // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

union V8 {
  uint4 u;
  uint16_t h[8];
  bf8 v;
};

// Sum of the 1x1 head over one group of 8 channels, in channel order,
// starting from s0.  The head adds the groups' sums in group order:
// out = ((b + S_0) + S_1) + ... with S_g the group sum started from 0 (g > 0),
// the order that the fused head of the convolution epilogue also follows.
__device__ __forceinline__ float head_group(const V8& v, const float* w, float s0) {
  float s = s0;
#pragma unroll
  for (int k = 0; k < 8; k++) s = fmaf(bf2f(v.h[k]), w[k], s);
  return s;
}


__device__ __forceinline__ void store1(uint16_t* p, float v) { *p = f2bf(v); }
__device__ __forceinline__ float load1(const uint16_t* p) { return bf2f(*p); }
__device__ __forceinline__ float load1(const float* p) { return *p; }
__device__ __forceinline__ void store1(float* p, float v) { *p = v; }

// ----------------------------------------------------------------- conv
// Block: 256 threads = 4 waves on a 4R x 16 pixel tile of one image and COB
// output channels.  Per 32-channel chunk of the input the block stages the
// tile with its one-pixel halo ((4R + 2) x 18 pixels) and the chunk's
// weights for all taps in LDS (64-byte rows with a bank swizzle, kLdsRow);
// wave w then multiplies the COB/16 weight tiles by tile rows Rw..Rw+R-1 (R
// MFMA column tiles of 16 pixels) for every tap, reading both operands from
// LDS.
//
// The weights are the MFMA's A operand and the pixels its B operand: lane l
// receives rows 4(l/16)..4(l/16)+3 of every weight tile's result for pixel
// l % 16.  The weight rows are staged in the order that makes those rows the
// COB/4 consecutive output channels (COB/4)(l/16)... (slot_channel), so each
// lane holds whole 16-byte runs of channels of its pixels: the epilogue adds
// the bias, applies the ReLU, rounds to BF16 and stores, pools (the pixel
// pair x, x+1 sits in neighbouring lanes) or evaluates the fused head
// straight from registers.
//
// The grid is persistent: a block walks the work items (tile, chunk) of its
// output-channel block, tiles blockIdx.x, blockIdx.x + gridDim.x, ..., and
// loads the next item's input and weights into registers while the MFMAs of
// the current one run; the weights are loaded once when the input has a
// single chunk.  The input pieces (8 channels of one halo pixel) are buffer
// loads over the tile's image, at per-thread offsets fixed for the launch
// plus one offset per tile; a piece outside the image (or the channels) is
// given an offset past the image's end, which the buffer load returns as 0.
constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;
// LDS rows of staged pixels and weights: 64 bytes (32 BF16), the 16-byte
// piece c of LDS row p stored at position c ^ swz(p), swz(p) = (p / 2) % 4.
// The operand reads and the staging writes (ds_read_b128, ds_write_b128)
// take 16 consecutive rows x pieces g = l / 16 per wavefront; measured on
// gfx950, plain 64- or 80-byte rows read at half the rate of consecutive
// memory (80-byte rows write at full rate), 64-byte rows with this swizzle
// read and write at full rate.  A staged halo row is kPitch = 24 LDS rows
// (18 used), so that swz of a pixel depends on its column only.
constexpr int kLdsRow = 32;
constexpr int kPitch = 24;
__device__ __forceinline__ int swz(int x) { return (x >> 1) & 3; }
constexpr unsigned kOutside = 0x80000000u;         // a byte offset past every image

// Output channel (within the block's COB) of result slot s = 16 j + m, m the
// row of weight tile j's 16 x 16 result: lane group m / 4 holds channels
// (COB/4)(m/4) ... (COB/4)(m/4) + COB/4 - 1, in the order 4 j + m % 4.
template <int COB>
__device__ __forceinline__ int slot_channel(int s) {
  return (COB / 4) * ((s & 15) >> 2) + 4 * (s >> 4) + (s & 3);
}

// Buffer resource over `bytes` bytes at p: loads at larger offsets return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t image_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// s_waitcnt vmcnt(0).  A load whose value is stored under a per-lane
// condition is waited for only on the branch that stores it; without this
// wait on every path the compiler makes the next writes of the load's
// registers (the MFMA operands' LDS reads) wait for all loads in flight,
// the next item's prefetch included.
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// two floats rounded to BF16 (nearest even, as f2bf) in one dword, lo first
// (one v_cvt_pk_bf16_f32)
typedef __attribute__((ext_vector_type(2))) float f2v;
typedef __attribute__((ext_vector_type(2))) __bf16 bf2v;
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){lo, hi}, bf2v));
}
// max(v, 0) as an integer max of the bits (one instruction; the float max
// is preceded by a NaN-quieting copy): equal for every non-NaN v
__device__ __forceinline__ float relu(float v) { return __int_as_float(max(__float_as_int(v), 0)); }
__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

template <typename TX, typename TY, int COB, int R, bool FIRST>
__global__ __launch_bounds__(256, COB == 32 && R == 4 ? 3 : 2) void k_conv(ConvArgs<TX, TY> a) {
  constexpr int CT = COB / 16;
  constexpr int CL = COB / 4;                        // consecutive output channels per lane
  constexpr int PW = sizeof(TX) / 2;                 // 16-byte loads per input piece
  // BF16 input: the next chunk's weights are prefetched into registers with
  // its input; FP32 input (no registers to spare): they are loaded when
  // stored, all pieces in flight at once
  // FIRST: the input is computed from the FP32 images (a single chunk whose
  // weights are staged once, no prefetch)
  constexpr bool kWPrefetch = PW == 1 && !FIRST;
  constexpr int kTH = 4 * R;       // tile rows: R per wave
  constexpr int kHY = kTH + 2;     // halo rows
  // FIRST: the BF16 image over the halo's own halo (kRY x kRX pixels), per
  // pixel its row of three taps (the pixel and the next two) in 8 bytes;
  // one more zero entry
  constexpr int kRY = FIRST ? kHY + 2 : 1, kRX = kHalo + 2, kRawPer = (kRY * kRX + 255) / 256;
  __shared__ __align__(16) uint16_t s_in[kHY * kPitch * kLdsRow];
  __shared__ __align__(16) uint16_t s_w[9 * COB * kLdsRow];
  __shared__ __align__(16) uint2 s_raw3[FIRST ? kRY * kRX + 1 : 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const int tiles_x = (a.W + kTile - 1) / kTile, tiles_y = (a.H + kTH - 1) / kTH;
  const int ntiles = a.N * tiles_x * tiles_y;
  const int nck = (a.Cin + 31) / 32;
  const int co0 = blockIdx.y * COB;
  const int cl0 = co0 + CL * g;   // this lane's first output channel
  const int taps = a.taps;
  int tt = blockIdx.x;
  if (tt >= ntiles) return;
  float bl[CL];
#pragma unroll
  for (int k = 0; k < CL; k++) bl[k] = a.bias && cl0 + k < a.Cout ? a.bias[cl0 + k] : 0.f;
  // the fused head's weights of this lane's channels (COB 32 only), loaded
  // here: a load in the epilogue would wait for the prefetch in flight
  constexpr int HW = COB == 32 ? CL : 1;
  float hw[HW], hb = 0.f;
  if constexpr (COB == 32) {
    if (a.head_out) {
#pragma unroll
      for (int k = 0; k < HW; k++) hw[k] = a.head_w[cl0 + k];
      hb = g == 0 ? a.head_b[0] : 0.f;
    }
  }
  // ---- staging.  One wave instruction moves 16 consecutive pixels (or
  // weight rows) x 4 pieces of 8 channels: lane l takes pixel / row l % 16
  // and piece (l / 16) % 4, so each 32 lanes write 16 consecutive LDS rows
  // x 2 pieces (distinct banks, kLdsRow) and the wave reads 1 KB of
  // consecutive memory.  Offsets that do not change during the launch are
  // computed here once; a piece outside the image or the tensor is given
  // the offset kOutside, which the buffer load returns as 0.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int lx = lane & 15, lc8 = 8 * ((lane >> 4) & 3);
  // Input, main part: halo rows 0..kHY-1 x halo columns 0..15; wave w takes
  // the rows w, w + 4, ... (kRounds rounds, the last one waves 0 and 1 only).
  // Byte offset of this lane's piece from halo column 0 of a row, in x and
  // in the half-size tensor up:
  constexpr int kRounds = (kHY + 3) / 4;
  const unsigned lo_x = (unsigned)((lx * a.ldx + lc8) * (int)sizeof(TX));
  const unsigned lo_u = (unsigned)((((lx + 1) >> 1) * a.ldu + lc8) * 2);
  // Input, extra part: halo columns 16 and 17 of every row, 8 kHY pieces in
  // kExR rounds; piece e = t + 256 round: row e / 8, column 16 + (e / 4) % 2,
  // piece e % 4.
  constexpr int kExR = (8 * kHY + 255) / 256;
  const int ex_x = kTile + ((tid >> 2) & 1), ex_c8 = 8 * (tid & 3);
  // LDS element offsets of this lane's pieces (main: round 0; round r is
  // 4 r halo rows further)
  const int la_in = (wave * kPitch + lx) * kLdsRow + 8 * (((lane >> 4) & 3) ^ swz(lx));
  const int la_ex = ((tid >> 3) * kPitch + ex_x) * kLdsRow + 8 * ((tid & 3) ^ swz(ex_x));   // round e: + 32 e halo rows
  // Weights: slot row r = tap * COB + slot; round k covers rows 64 k ... 64 k
  // + 63, wave w the rows 64 k + 16 w + (0..15), so the tap (64 k + 16 w) /
  // COB of a round is uniform and the lane's slot (16 w + lx) % COB fixed.
  constexpr int kWRounds = (9 * COB + 63) / 64;
  const int wco = co0 + slot_channel<COB>((16 * wave + lx) % COB);
  const unsigned lo_w = wco < a.Cout ? (unsigned)((wco * a.Cin + lc8) * 2) : kOutside;
  const int la_w = (16 * wave + lx) * kLdsRow + 8 * (((lane >> 4) & 3) ^ swz(lx));
  uint4 pin[FIRST ? 1 : kRounds + kExR][PW];
  uint4 wreg[kWRounds];
  // loads the chunk's weights into registers
  const uint16_t* const Wt0 = a.w;
  const long wtap = (long)a.Cout * a.Cin;   // elements per tap
  auto wfetch = [&](int ck) {
    const uint16_t* wck = Wt0 + 32 * ck;
#pragma unroll
    for (int k = 0; k < kWRounds; k++) {
      const int t = (64 * k + 16 * wv) / COB;
      wreg[k] = make_uint4(0, 0, 0, 0);
      if (t < taps) {
        const auto rs = image_rsrc(wck + t * wtap, 0x7fffffffu);
        wreg[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, lo_w, 0, 0));
      }
    }
  };
  auto wput = [&]() {
#pragma unroll
    for (int k = 0; k < kWRounds; k++)
      if ((64 * k + 16 * wv) / COB < taps) *reinterpret_cast<uint4*>(&s_w[la_w + 64 * k * kLdsRow]) = wreg[k];
    vm_wait_all();
  };
  // loads the input tile + halo of work item (tile, chunk) into registers
  const TX* const X = a.x;
  const uint16_t* const UP = a.up;
  const int H = a.H, W = a.W, H2 = a.H >> 1, W2 = a.W >> 1, ldx = a.ldx, ldu = a.ldu;
  // FIRST: this thread's image pixels p = tid + 256 i of the kRY x kRX
  // region at image (y0 - 2, x0 - 2) and the two after each; the first
  // convolution's weights as MFMA A operands, K position 4 ty + tx for tap
  // (ty, tx) (positions 4 ty + 3 and from 12 on zero), and its bias as the
  // MFMA's C operand, per lane as the second convolution's own slots
  float rawv[FIRST ? kRawPer : 1][3];
  bf8 w1a[FIRST ? CT : 1];
  f4 b1c[FIRST ? CT : 1];
  if constexpr (FIRST) {
#pragma unroll
    for (int j = 0; j < CT; j++) {
      const int ch = slot_channel<COB>(16 * j + r16);
      V8 v;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int ty = 2 * g + k / 4, tx = k % 4;
        v.h[k] = tx < 3 && ty < 3 ? a.w1[((long)(3 * ty + tx) * COB + ch) * a.w1_cin] : (uint16_t)0;
      }
      w1a[j] = v.v;
      // result rows 4 g + r of weight tile j: channels cl0 + 4 j + r
      b1c[j] = f4{a.b1[cl0 + 4 * j], a.b1[cl0 + 4 * j + 1], a.b1[cl0 + 4 * j + 2], a.b1[cl0 + 4 * j + 3]};
    }
    if (tid == 0) s_raw3[kRY * kRX] = make_uint2(0, 0);
  }
  auto fetch = [&](int tile, int ck) {
    const int tx = tile % tiles_x, rest = tile / tiles_x;
    const int y0 = (rest % tiles_y) * kTH, x0 = tx * kTile, n = rest / tiles_y;
    if constexpr (FIRST) {
      const auto rs = image_rsrc(a.x1 + (long)n * a.H * a.W, 0x7fffffffu);
#pragma unroll
      for (int i = 0; i < kRawPer; i++) {
        const int p = tid + 256 * i, ry = p / kRX, rx = p - kRX * ry;
        const int yy = y0 - 2 + ry, xx = x0 - 2 + rx;
        const bool oky = p < kRY * kRX && (unsigned)yy < (unsigned)a.H;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const bool ok = oky && (unsigned)(xx + d) < (unsigned)a.W;
          rawv[i][d] = __uint_as_float(
              __builtin_amdgcn_raw_buffer_load_b32(rs, ok ? (unsigned)((yy * a.W + xx + d) * 4) : kOutside, 0, 0));
        }
      }
      return;
    }
    const int ci0 = 32 * ck;
    bool from_up = false;
    if constexpr (sizeof(TX) == 2) from_up = ci0 < a.Cup;   // Cup is a multiple of 32
    // halo row wv (image row y0 - 1 + wv) from halo column 0 (image column
    // x0 - 1; its offset is not used where that column is outside the
    // image), and the step of 4 halo rows (2 rows of up), in bytes
    const int yw = y0 - 1 + wv;
    const char* row0;
    long step;
    if (from_up) {
      row0 = reinterpret_cast<const char*>(UP + (((long)n * H2 + (yw >> 1)) * W2 + (x0 >> 1) - 1) * ldu + a.cu0 + ci0);
      step = 2L * W2 * ldu * 2;
    } else {
      row0 = reinterpret_cast<const char*>(X + (((long)n * H + yw) * W + x0 - 1) * ldx + a.cx0 + ci0);
      step = 4L * W * ldx * (long)sizeof(TX);
    }
    const unsigned lo = (unsigned)(x0 - 1 + lx) < (unsigned)W ? (from_up ? lo_u : lo_x) : kOutside;
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
#pragma unroll
      for (int h = 0; h < PW; h++) pin[r][h] = make_uint4(0, 0, 0, 0);
      if ((r < kRounds - 1 || wv + 4 * r < kHY) && (unsigned)(yw + 4 * r) < (unsigned)H) {
        const auto rs = image_rsrc(row0 + r * step, 0x7fffffffu);
#pragma unroll
        for (int h = 0; h < PW; h++)
          pin[r][h] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, lo + 16 * h, 0, 0));
      }
    }
#pragma unroll
    for (int e = 0; e < kExR; e++) {
      const int yy = y0 - 1 + (tid >> 3) + 32 * e, xx = x0 - 1 + ex_x;
      const bool ok = tid + 256 * e < 8 * kHY && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
      const char* img;
      unsigned off;
      if (from_up) {
        img = reinterpret_cast<const char*>(UP + (long)n * H2 * W2 * ldu + a.cu0 + ci0);
        off = (unsigned)((((yy >> 1) * W2 + (xx >> 1)) * ldu + ex_c8) * 2);
      } else {
        img = reinterpret_cast<const char*>(X + (long)n * H * W * ldx + a.cx0 + ci0);
        off = (unsigned)(((yy * W + xx) * ldx + ex_c8) * (int)sizeof(TX));
      }
      const auto rs = image_rsrc(img, 0x7fffffffu);
#pragma unroll
      for (int h = 0; h < PW; h++)
        pin[kRounds + e][h] =
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, ok ? off + 16 * h : kOutside, 0, 0));
    }
  };
  // FIRST: the staged input is the first convolution over the halo, from the
  // image rows in LDS: per 16 halo pixels one MFMA per 16 channels, K = the
  // taps (lane group 0 tap rows 0 and 1, group 1 row 2, the rest zero); zero
  // outside the image (the input's padding); the tile's own pixels also to
  // a1_out
  auto put_first = [&](int tile) {
    const int tx = tile % tiles_x, rest = tile / tiles_x;
    const int y0 = (rest % tiles_y) * kTH, x0 = tx * kTile, n = rest / tiles_y;
    const float sc = a.x1_scale ? a.x1_scale[n] : 1.f;
#pragma unroll
    for (int i = 0; i < kRawPer; i++) {
      const int p = tid + 256 * i;
      // as the input conversion rounds the image
      if (p < kRY * kRX)
        s_raw3[p] = make_uint2((uint32_t)f2bf(rawv[i][0] / sc) | ((uint32_t)f2bf(rawv[i][1] / sc) << 16),
                               (uint32_t)f2bf(rawv[i][2] / sc));
    }
    __syncthreads();
    constexpr int kHP = kHY * kHalo, kPT = (kHP + 15) / 16, kZero = kRY * kRX;
    // every halo pixel inside the image (most tiles): no per-pixel test
    const bool interior = y0 >= 1 && y0 + kTH + 1 <= a.H && x0 >= 1 && x0 + kTile + 1 <= a.W;
    // halo pixel q = 16 pt + r16, pt = wv, wv + 4, ...: q grows by 64 = 3 rows + 10 columns
    int hy = (16 * wv + r16) / kHalo, hx = (16 * wv + r16) - kHalo * hy;
#pragma unroll 1
    for (int pt = wv; pt < kPT; pt += 4) {
      const bool valid = 16 * pt + r16 < kHP;
      const int rb = hy * kRX + hx;   // image pixel of tap (0, 0)
      const int i0 = !valid ? kZero : g == 0 ? rb : g == 1 ? rb + 2 * kRX : kZero;
      const int i1 = valid && g == 0 ? rb + kRX : kZero;
      const uint2 lo = s_raw3[i0], hi = s_raw3[i1];
      V8 b;
      b.u = make_uint4(lo.x, lo.y, hi.x, hi.y);
      f4 d[CT];
#pragma unroll
      for (int j = 0; j < CT; j++) d[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1a[j], b.v, b1c[j], 0, 0, 0);
      // channels cl0 + k = 8 g + k of halo pixel (hy, hx)
      const int yy = y0 - 1 + hy, xx = x0 - 1 + hx;
      const bool inside = valid && (interior || ((unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W));
      uint4 o;
      uint32_t* po = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
      for (int k = 0; k < CL; k += 2)
        po[k / 2] = pack_bf16(relu(d[k >> 2][k & 3]), relu(d[(k + 1) >> 2][(k + 1) & 3]));
      if (!inside) o = make_uint4(0, 0, 0, 0);
      if (valid) *reinterpret_cast<uint4*>(&s_in[(hy * kPitch + hx) * kLdsRow + 8 * (g ^ swz(hx))]) = o;
      if (a.a1_out && inside && hy >= 1 && hy <= kTH && hx >= 1 && hx <= kTile)
        *reinterpret_cast<uint4*>(a.a1_out + (((long)n * a.H + yy) * a.W + xx) * COB + cl0) = o;
      hx += 10;
      hy += 3;
      if (hx >= kHalo) {
        hx -= kHalo;
        hy += 1;
      }
    }
  };
  // stores the prefetched input and, when they change, the chunk's weights
  // (read straight from L2: every block of the launch reads the same ones)
  auto put = [&](int tile, int ck, bool with_w) {
    auto bf = [&](const uint4* p) {
      if constexpr (PW == 1) {
        return p[0];
      } else {
        // FP32 input: 8 channels rounded to BF16
        uint4 v;
        v.x = pack_bf16(__uint_as_float(p[0].x), __uint_as_float(p[0].y));
        v.y = pack_bf16(__uint_as_float(p[0].z), __uint_as_float(p[0].w));
        v.z = pack_bf16(__uint_as_float(p[1].x), __uint_as_float(p[1].y));
        v.w = pack_bf16(__uint_as_float(p[1].z), __uint_as_float(p[1].w));
        return v;
      }
    };
    if constexpr (FIRST) {
      put_first(tile);
    } else {
#pragma unroll
      for (int r = 0; r < kRounds; r++)
        if (r < kRounds - 1 || wv + 4 * r < kHY)
          *reinterpret_cast<uint4*>(&s_in[la_in + 4 * r * kPitch * kLdsRow]) = bf(pin[r]);
#pragma unroll
      for (int e = 0; e < kExR; e++)
        if (tid + 256 * e < 8 * kHY)
          *reinterpret_cast<uint4*>(&s_in[la_ex + 32 * e * kPitch * kLdsRow]) = bf(pin[kRounds + e]);
      vm_wait_all();
    }
    if (with_w) {
      if constexpr (!kWPrefetch) wfetch(ck);
      wput();
    }
  };
  f4 acc[R][CT];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int j = 0; j < CT; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  int ck = 0;
  bool w_in = true;   // the weights in LDS are not this chunk's
  fetch(tt, 0);
  if constexpr (kWPrefetch) wfetch(0);
  for (;;) {
    put(tt, ck, w_in);
    __syncthreads();
    // the next work item; its loads are in flight during this item's MFMAs
    int tt2 = tt, ck2 = ck + 1;
    if (ck2 == nck) {
      ck2 = 0;
      tt2 = tt + gridDim.x;
    }
    const bool more = tt2 < ntiles;
    w_in = nck > 1;
    if (more) {
      fetch(tt2, ck2);
      if constexpr (kWPrefetch)
        if (w_in) wfetch(ck2);
    }
    if (taps == 9) {
      // for a column offset dx the B operand of tile row i at row offset dy
      // is halo row R w + i + dy: R + 2 halo rows serve all three dy
#pragma unroll 1
      for (int dx = 0; dx < 3; dx++) {
        // scheduling hint: interleave the LDS operand reads with the MFMAs
        // (single-wave small-GEMM strategy; measured 1.7 % faster, same
        // code otherwise)
        __builtin_amdgcn_iglp_opt(1);
        bf8 P[R + 2];
        const int pc = 8 * (g ^ swz(r16 + dx));
#pragma unroll
        for (int h = 0; h < R + 2; h++) {
          V8 v;
          v.u = *reinterpret_cast<const uint4*>(&s_in[((R * wave + h) * kPitch + r16 + dx) * kLdsRow + pc]);
          P[h] = v.v;
        }
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
          const int t = dy * 3 + dx;
          bf8 Wt[CT];
#pragma unroll
          for (int j = 0; j < CT; j++) {
            V8 v;
            v.u = *reinterpret_cast<const uint4*>(&s_w[(t * COB + 16 * j + r16) * kLdsRow + 8 * (g ^ swz(r16))]);
            Wt[j] = v.v;
          }
#pragma unroll
          for (int i = 0; i < R; i++)
#pragma unroll
            for (int j = 0; j < CT; j++)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wt[j], P[i + dy], acc[i][j], 0, 0, 0);
        }
      }
    } else {
      // 1x1: the centre of the halo
      bf8 Wt[CT];
#pragma unroll
      for (int j = 0; j < CT; j++) {
        V8 v;
        v.u = *reinterpret_cast<const uint4*>(&s_w[(16 * j + r16) * kLdsRow + 8 * (g ^ swz(r16))]);
        Wt[j] = v.v;
      }
#pragma unroll
      for (int i = 0; i < R; i++) {
        V8 v;
        v.u = *reinterpret_cast<const uint4*>(&s_in[((R * wave + i + 1) * kPitch + r16 + 1) * kLdsRow + 8 * (g ^ swz(r16 + 1))]);
#pragma unroll
        for (int j = 0; j < CT; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Wt[j], v.v, acc[i][j], 0, 0, 0);
      }
    }
    if (ck == nck - 1) {
      // epilogue: acc[i][k / 4][k % 4] is channel cl0 + k of pixel
      // (y0 + R wave + i, x0 + r16)
      const int tx = tt % tiles_x, rest = tt / tiles_x;
      const int y0 = (rest % tiles_y) * kTH, x0 = tx * kTile, n = rest / tiles_y;
      const int x = x0 + r16;
      if constexpr (sizeof(TY) == 2) {
#pragma unroll
        for (int i0 = 0; i0 < R; i0 += 2) {
          // rows i0, i0 + 1: bias, ReLU and the rounding to BF16 once (the
          // stores, the pool and the head all use the rounded values)
          uint32_t pk[2][CL / 2];
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const int i = i0 + h;
#pragma unroll
            for (int k = 0; k < CL; k += 2) {
              float v0 = acc[i][k >> 2][k & 3] + bl[k];
              float v1 = acc[i][(k + 1) >> 2][(k + 1) & 3] + bl[k + 1];
              if (a.relu) {
                v0 = relu(v0);
                v1 = relu(v1);
              }
              pk[h][k / 2] = pack_bf16(v0, v1);
            }
            const int y = y0 + R * wave + i;
            if (COB == 32 && a.head_out) {
              // lane group g holds the head's channel group g (COB = Cout
              // = 32): its sum in channel order, then the groups' sums in
              // group order, as k_head_fwd adds them
              float s = hb;
#pragma unroll
              for (int k = 0; k < HW; k += 2) {
                s = fmaf(bf_lo(pk[h][k / 2]), hw[k], s);
                s = fmaf(bf_hi(pk[h][k / 2]), hw[k + 1], s);
              }
              const float s1 = __shfl(s, r16 + 16), s2 = __shfl(s, r16 + 32), s3 = __shfl(s, r16 + 48);
              if (g == 0 && y < a.H && x < a.W) a.head_out[((long)n * a.H + y) * a.W + x] = ((s + s1) + s2) + s3;
            } else if (y < a.H && x < a.W) {
              uint16_t* dst = a.y + (((long)n * a.H + y) * a.W + x) * a.ldy + a.cy0 + cl0;
#pragma unroll
              for (int h8 = 0; h8 < CL / 8; h8++)
                if (cl0 + 8 * h8 < a.Cout)
                  *reinterpret_cast<uint4*>(dst + 8 * h8) =
                      make_uint4(pk[h][4 * h8], pk[h][4 * h8 + 1], pk[h][4 * h8 + 2], pk[h][4 * h8 + 3]);
            }
          }
          if (a.pool) {
            // 2x2 average of the rounded values, summed in the order
            // (y, x), (y, x+1), (y+1, x), (y+1, x+1); pixel x + 1 is the
            // next lane's (DPP row_shl:1 reads lane l + 1 of the 16-lane row)
            uint32_t nb[2][CL / 2];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
              for (int k = 0; k < CL / 2; k++)
                nb[h][k] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pk[h][k], 0x101, 0xf, 0xf, true);
            const int y = y0 + R * wave + i0;
            if (!(r16 & 1) && y + 1 < a.H && x + 1 < a.W) {
              uint32_t po[CL / 2];
#pragma unroll
              for (int k = 0; k < CL / 2; k++) {
                const float lo = 0.25f * (((bf_lo(pk[0][k]) + bf_lo(nb[0][k])) + bf_lo(pk[1][k])) + bf_lo(nb[1][k]));
                const float hi = 0.25f * (((bf_hi(pk[0][k]) + bf_hi(nb[0][k])) + bf_hi(pk[1][k])) + bf_hi(nb[1][k]));
                po[k] = pack_bf16(lo, hi);
              }
              uint16_t* dst =
                  a.pool + (((long)n * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1)) * a.ldp + a.cp0 + cl0;
#pragma unroll
              for (int h8 = 0; h8 < CL / 8; h8++)
                if (cl0 + 8 * h8 < a.Cout)
                  *reinterpret_cast<uint4*>(dst + 8 * h8) =
                      make_uint4(po[4 * h8], po[4 * h8 + 1], po[4 * h8 + 2], po[4 * h8 + 3]);
            }
          }
        }
      } else {
        // FP32 output (masked, accumulating), one element at a time
#pragma unroll
        for (int i = 0; i < R; i++) {
          const int y = y0 + R * wave + i;
          if (y >= a.H || x >= a.W) continue;
          const long p = ((long)n * a.H + y) * a.W + x;
#pragma unroll
          for (int k = 0; k < CL; k++) {
            const int co = cl0 + k;
            if (co >= a.Cout) continue;
            float v = acc[i][k >> 2][k & 3] + bl[k];
            if (a.relu) v = relu(v);
            if (a.mask && !(bf2f(a.mask[p * a.ldm + a.cm0 + co]) > 0.f)) v = 0.f;
            TY* dst = a.y + p * a.ldy + a.cy0 + co;
            if (a.accumulate) v += load1(dst);
            store1(dst, v);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < R; i++)
#pragma unroll
        for (int j = 0; j < CT; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();   // the LDS operands are consumed before the next put
    if (!more) break;
    tt = tt2;
    ck = ck2;
  }
}

// ------------------------------------------------------------ weight grad
// dW[tap][co][ci] += sum_p dY[p][co] * X[p + off(tap)][ci] over a pixel chunk.
// Block 256 threads: output tile 64 co x 64 ci for one tap; K = pixels in
// steps of 32 staged (transposed) in LDS; split-K over blockIdx.z, each split
// writing its sums to part[z] (added to dW in split order afterwards).
__global__ __launch_bounds__(256) void k_conv_wgrad(WgradArgs a) {
  __shared__ uint16_t sdy[64][40];   // [co][pixel], padded rows
  __shared__ uint16_t sx[64][40];    // [ci][pixel]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tap = blockIdx.x % a.taps;
  const int co_t = (blockIdx.x / a.taps) % ((a.Cout + 63) / 64);
  const int ci_t = blockIdx.x / a.taps / ((a.Cout + 63) / 64);
  const int half = a.taps == 9 ? 1 : 0;
  const int dy = tap / 3 * half - half, dx = tap % 3 * half - half;
  const long P = (long)a.N * a.H * a.W;
  const long chunk = (P + gridDim.z - 1) / gridDim.z;
  const long p0 = (long)blockIdx.z * chunk, p1 = std::min(P, p0 + chunk);
  const int r16 = lane & 15, g = lane >> 4;
  const int co_w = 32 * (wave >> 1), ci_w = 32 * (wave & 1);
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // staging: thread t loads pixel (t & 31), 8 channels at 8*(t >> 5)
  const int sp = threadIdx.x & 31, sc = 8 * (threadIdx.x >> 5);
  for (long pb = p0; pb < p1; pb += 32) {
    const long p = pb + sp;
    {
      // dY[p][co_t*64 + sc ..]
      const int co = co_t * 64 + sc;
      float v[8];
      const bool ok = p < p1 && co < a.Cout;
      if (ok) {
        const float4 q0 = *reinterpret_cast<const float4*>(a.dy + p * a.lddy + a.cdy0 + co);
        const float4 q1 = *reinterpret_cast<const float4*>(a.dy + p * a.lddy + a.cdy0 + co + 4);
        v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w; v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
      }
#pragma unroll
      for (int q = 0; q < 8; q++) sdy[sc + q][sp] = ok ? f2bf(v[q]) : (uint16_t)0;
      // X[p + off][ci_t*64 + sc ..]
      const int ci = ci_t * 64 + sc;
      bool okx = p < p1 && ci < a.Cin;
      long src = 0;
      if (okx) {
        const int xx = (int)(p % a.W), yy = (int)((p / a.W) % a.H);
        const long n = p / ((long)a.W * a.H);
        const int y2 = yy + dy, x2 = xx + dx;
        okx = y2 >= 0 && y2 < a.H && x2 >= 0 && x2 < a.W;
        src = ((n * a.H + y2) * a.W + x2) * a.ldx + a.cx0 + ci;
      }
      V8 xv;
      xv.u = okx ? *reinterpret_cast<const uint4*>(a.x + src) : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 8; q++) sx[sc + q][sp] = xv.h[q];
    }
    __syncthreads();
    // A[co][k = pixel] from sdy, B[k = pixel][ci] from sx
    bf8 A[2], B[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      V8 t;
      t.u = *reinterpret_cast<const uint4*>(&sdy[co_w + 16 * i + r16][8 * g]);
      A[i] = t.v;
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      V8 t;
      t.u = *reinterpret_cast<const uint4*>(&sx[ci_w + 16 * j + r16][8 * g]);
      B[j] = t.v;
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D[co = 4g + r][ci = r16]
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int co = co_t * 64 + co_w + 16 * i + 4 * g + r;
        const int ci = ci_t * 64 + ci_w + 16 * j + r16;
        if (co < a.Cout && ci < a.Cin)
          a.part[(((long)blockIdx.z * a.taps + tap) * a.Cout + co) * a.Cin + ci] = acc[i][j][r];
      }
}

// ------------------------------------------------------------ small ops

// parts[blockIdx.y][c] = this block's share of sum_p dY[p][c]
__global__ __launch_bounds__(256) void k_bias_grad(const float* dy, int ld, int c0, long P, int C, float* parts) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int part = threadIdx.x >> 6;
  __shared__ float red[4][64];
  float s = 0.f;
  if (c < C)
    for (long p = blockIdx.y * 4 + part; p < P; p += 4L * gridDim.y) s += dy[p * ld + c0 + c];
  red[part][threadIdx.x & 63] = s;
  __syncthreads();
  if (part == 0 && c < C)
    parts[(long)blockIdx.y * C + c] = (red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// 2x2 average pool (bf16 NHWC, channel sub-ranges); 8 channels per thread
// (C, strides and channel offsets are multiples of 8).
__global__ void k_pool_bwd(PoolGradArgs a) {
  const long n = (long)a.N * a.H * a.W * a.C;   // H, W of the input
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % a.C);
    const long q = i / a.C;
    const int x = (int)(q % a.W), y = (int)((q / a.W) % a.H);
    const long nn = q / ((long)a.W * a.H);
    const float v = 0.25f * a.dout[((nn * (a.H / 2) + y / 2) * (a.W / 2) + x / 2) * a.ldo + a.co0 + c];
    float* d = a.din + q * a.ldi + a.ci0 + c;
    *d = a.accumulate ? *d + v : v;
  }
}

// nearest 2x upsample (bf16): out[2y+u][2x+v] = in[y][x]; 8 channels per thread
__global__ void k_up_fwd(PoolArgs a) {
  const int C8 = a.C / 8;
  const long n = (long)a.N * (2 * a.H) * (2 * a.W) * C8;   // H, W of the input
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = 8 * (int)(i % C8);
    const long q = i / C8;
    const int x = (int)(q % (2 * a.W)), y = (int)((q / (2 * a.W)) % (2 * a.H));
    const long nn = q / ((long)(2 * a.W) * (2 * a.H));
    *reinterpret_cast<uint4*>(a.out + q * a.ldo + a.co0 + c) =
        *reinterpret_cast<const uint4*>(a.in + ((nn * a.H + y / 2) * a.W + x / 2) * a.ldi + a.ci0 + c);
  }
}

// gradient of the upsample: dIn[y][x] = sum of the 2x2 dOut block (fp32)
__global__ void k_up_bwd(PoolGradArgs a) {
  const long n = (long)a.N * a.H * a.W * a.C;   // H, W of the (small) input
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % a.C);
    const long q = i / a.C;
    const int x = (int)(q % a.W), y = (int)((q / a.W) % a.H);
    const long nn = q / ((long)a.W * a.H);
    const long b = ((nn * 2 * a.H + 2 * y) * 2 * a.W + 2 * x) * a.ldo + a.co0 + c;
    const float v = a.dout[b] + a.dout[b + a.ldo] + a.dout[b + 2L * a.W * a.ldo] + a.dout[b + 2L * a.W * a.ldo + a.ldo];
    float* d = a.din + q * a.ldi + a.ci0 + c;
    *d = a.accumulate ? *d + v : v;
  }
}

// dIn *= (act > 0) in place (fp32 gradient, bf16 activation)
__global__ void k_relu_mask(float* g, int ldg, int cg0, const uint16_t* act, int lda, int ca0, long P, int C) {
  const long n = P * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / C;
    const int c = (int)(i % C);
    if (!(bf2f(act[p * lda + ca0 + c]) > 0.f)) g[p * ldg + cg0 + c] = 0.f;
  }
}

// 1x1 head: out[p] = b + sum_c x[p][c] w[c]  (fp32 out); C a multiple of 8
__global__ void k_head_fwd(const uint16_t* x, int ldx, const float* w, const float* b, int C, long P, float* out) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < C; c += 8) {
      V8 v;
      v.u = *reinterpret_cast<const uint4*>(x + p * ldx + c);
      const float g = head_group(v, w + c, c == 0 ? b[0] : 0.f);
      s = c == 0 ? g : s + g;
    }
    out[p] = s;
  }
}

// head backward: dX[p][c] = dout[p] w[c] (masked by ReLU of x); each block's
// sums of dw and db to part[block] (added in block order afterwards)
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t* x, int ldx, const float* w, int C, long P,
                                                 const float* dout, float* dx, int lddx, float* part) {
  __shared__ float red[256];
  float gw[64];
  for (int c = 0; c < C; c++) gw[c] = 0.f;
  float gb = 0.f;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    const float d = dout[p];
    gb += d;
    for (int c = 0; c < C; c++) {
      const float xv = bf2f(x[p * ldx + c]);
      gw[c] = fmaf(d, xv, gw[c]);
      dx[p * lddx + c] = xv > 0.f ? d * w[c] : 0.f;
    }
  }
  for (int c = 0; c <= C; c++) {
    red[threadIdx.x] = c < C ? gw[c] : gb;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * (C + 1) + c] = red[0];   // [block][dw..., db]
    __syncthreads();
  }
}

// dgrad weights: wt[tap'][ci][co] = w[8 - tap'][co][ci]  (bf16)
__global__ void k_flip_weights(const uint16_t* w, uint16_t* wt, int taps, int Cout, int Cin) {
  const long n = (long)taps * Cout * Cin;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int co = (int)(i % Cout);
    const int ci = (int)((i / Cout) % Cin);
    const int tp = (int)(i / ((long)Cout * Cin));
    wt[i] = w[((long)(taps - 1 - tp) * Cout + co) * Cin + ci];
  }
}

// Adam step on fp32 master weights, refreshes the bf16 copy.
__global__ void k_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
  }
}

__global__ void k_f32_to_bf16(const float* in, uint16_t* out, long n, int ldo, int co0, int C) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long p = i / C;
    const int c = (int)(i % C);
    out[p * ldo + co0 + c] = f2bf(in[i]);
  }
}

// MSE loss on scaled residual targets: d = pred - target; loss partial sums
__global__ __launch_bounds__(256) void k_mse(const float* pred, const float* target, long P, float* dpred,
                                            double* loss, float scale) {
  __shared__ double red[256];
  double s = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    const float d = pred[p] - target[p];
    s += (double)d * d;
    dpred[p] = scale * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Per-image scale of the warm start: s = RMS of the image (1 for a zero image),
// accumulated in FP64: kRmsParts blocks per image write partial sums, then one
// thread per image adds them in a fixed order.
constexpr int kRmsParts = 64;

__global__ __launch_bounds__(256) void k_image_rms_part(const float* in, long P, double* part) {
  __shared__ double red[256];
  const float* x = in + (long)blockIdx.y * P;
  double s = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x)
    s += (double)x[p] * x[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)blockIdx.y * kRmsParts + blockIdx.x] = red[0];
}

__global__ void k_image_rms_final(const double* part, long P, int B, float* scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s = 0.0;
  for (int k = 0; k < kRmsParts; k++) s += part[(long)b * kRmsParts + k];
  const double r = sqrt(s / (double)P);
  scale[b] = r > 0.0 ? (float)r : 1.f;
}

// out[p][0] = bf16(in[p] / scale[image]) into an 8-channel NHWC buffer
__global__ void k_to_bf16_scaled(const float* in, const float* scale, long P, long n, uint16_t* out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i * 8] = f2bf(in[i] / scale[i / P]);
}

// out = m + scale[image] * y
__global__ void k_residual_add(const float* m, const float* y, const float* scale, long P, long n, float gain,
                               float* out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = gain * fmaf(scale[i / P], y[i], m[i]);
}

__global__ void k_zero(float* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.f;
}

// ================================================================ launchers

static int blocks_for(long n, int per = 256) { return (int)std::min<long>((n + per - 1) / per, 65536); }

// Rows per wave of the 32-channel convolutions with BF16 input (tiles of
// 4 kRows32 x 16 pixels); the others use 4 (their registers are full).
constexpr int kRows32 = 8;

template <typename TX, typename TY, int COB, int R, bool FIRST = false>
void launch_conv_tiles(const ConvArgs<TX, TY>& a, int cus, hipStream_t st) {
  // LDS per block: the staged halo tile and one chunk's weights
  constexpr long lds = ((4L * R + 2) * kPitch + 9L * COB) * kLdsRow * 2 + (FIRST ? ((4L * R + 4) * (kHalo + 2) + 1) * 8 : 0);
  constexpr long per_cu = std::min<long>(COB == 32 && R == 4 ? 3 : 2, (160L * 1024) / lds);
  const long tiles = (long)a.N * ((a.H + 4 * R - 1) / (4 * R)) * ((a.W + kTile - 1) / kTile);
  const int cb = (a.Cout + COB - 1) / COB;
  const long gx = std::max<long>(1, std::min<long>(tiles, per_cu * cus / cb));
  hipLaunchKernelGGL((k_conv<TX, TY, COB, R, FIRST>), dim3((unsigned)gx, (unsigned)cb), dim3(256), 0, st, a);
}

template <typename TX, typename TY>
hipError_t launch_conv(const ConvArgs<TX, TY>& a, hipStream_t st) {
  if (a.taps != 9 && a.taps != 1) return hipErrorInvalidValue;
  // the fused head needs all 32 output channels in one block's BF16 epilogue
  if (a.head_out && (sizeof(TY) != 2 || a.Cout != 32 || a.mask || a.accumulate || a.pool || a.ldy % 8 ||
                     a.cy0 % 8 || !a.head_w || !a.head_b))
    return hipErrorInvalidValue;
  // BF16 output is plain (no mask, no accumulation) and stored 8 channels at
  // a time, as is its optional pool; FP32 output has no pool
  if (sizeof(TY) == 2 && (a.mask || a.accumulate || a.Cout % 8 || a.ldy % 8 || a.cy0 % 8)) return hipErrorInvalidValue;
  if (a.pool && (sizeof(TY) != 2 || a.ldp % 8 || a.cp0 % 8)) return hipErrorInvalidValue;
  // the input is staged in whole 32-channel chunks, each from x or from up;
  // one image's bytes are addressed by 32-bit buffer offsets
  if (a.Cin % 32 || (a.up && (a.Cup % 32 || a.H % 2 || a.W % 2)) || (!a.up && a.Cup) ||
      (long)a.H * a.W * a.ldx * (long)sizeof(TX) >= (1L << 31) || (long)a.H * a.W * a.ldu * 2 >= (1L << 31))
    return hipErrorInvalidValue;
  if (sizeof(TX) != 2 && a.up) return hipErrorInvalidValue;
  // persistent grid: the blocks that fit the device at once (LDS-bound), over
  // all output-channel blocks
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
  }
  if (a.x1) {
    // the input computed from the images: one 32-channel chunk, plain 3x3
    if (sizeof(TX) != 2 || sizeof(TY) != 2 || a.Cin != 32 || a.Cout != 32 || a.taps != 9 || a.up || !a.w1 ||
        !a.b1 || a.w1_cin < 1 || (long)a.N * a.H * a.W * 4 >= (1L << 31))
      return hipErrorInvalidValue;
    if constexpr (sizeof(TX) == 2 && sizeof(TY) == 2) launch_conv_tiles<TX, TY, 32, kRows32, true>(a, cus, st);
  } else if (a.Cout <= 32 || (sizeof(TX) == 2 && a.Cin >= 128)) {
    // (BF16 layers with 4 or more input chunks: measured faster on 32-channel
    // 32 x 16 tiles than on 64-channel 16 x 16 ones)
    launch_conv_tiles<TX, TY, 32, sizeof(TX) == 2 ? kRows32 : 4>(a, cus, st);
  } else {
    launch_conv_tiles<TX, TY, 64, 4>(a, cus, st);
  }
  return hipGetLastError();
}
template hipError_t launch_conv<uint16_t, uint16_t>(const ConvArgs<uint16_t, uint16_t>&, hipStream_t);
template hipError_t launch_conv<float, float>(const ConvArgs<float, float>&, hipStream_t);

hipError_t launch_wgrad(const WgradArgs& a, int splits, hipStream_t st) {
  if (!a.part || splits < 1) return hipErrorInvalidValue;
  const int nco = (a.Cout + 63) / 64, nci = (a.Cin + 63) / 64;
  hipLaunchKernelGGL(k_conv_wgrad, dim3(a.taps * nco * nci, 1, splits), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum_parts(a.part, splits, (long)a.taps * a.Cout * a.Cin, a.dw, st);
}

hipError_t launch_bias_grad(const float* dy, int ld, int c0, long P, int C, float* db, float* part, hipStream_t st) {
  hipLaunchKernelGGL(k_bias_grad, dim3((C + 63) / 64, kBiasGradParts), dim3(256), 0, st, dy, ld, c0, P, C, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum_parts(part, kBiasGradParts, C, db, st);
}

__global__ void k_sum_parts(const float* __restrict__ part, int nparts, long n, float* out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float s = part[i];
    for (int z = 1; z < nparts; z++) s += part[(long)z * n + i];
    out[i] += s;
  }
}

hipError_t launch_sum_parts(const float* part, int nparts, long n, float* out, hipStream_t st) {
  if (n <= 0 || nparts <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sum_parts, dim3(blocks_for(n)), dim3(256), 0, st, part, nparts, n, out);
  return hipGetLastError();
}

hipError_t launch_pool_bwd(const PoolGradArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for((long)a.N * a.H * a.W * a.C)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_up_fwd(const PoolArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_up_fwd, dim3(blocks_for((long)a.N * a.H * a.W * a.C / 2)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_up_bwd(const PoolGradArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for((long)a.N * a.H * a.W * a.C)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_relu_mask(float* g, int ldg, int cg0, const uint16_t* act, int lda, int ca0, long P, int C,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(P * C)), dim3(256), 0, st, g, ldg, cg0, act, lda, ca0, P, C);
  return hipGetLastError();
}
hipError_t launch_head_fwd(const uint16_t* x, int ldx, const float* w, const float* b, int C, long P, float* out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(P)), dim3(256), 0, st, x, ldx, w, b, C, P, out);
  return hipGetLastError();
}
hipError_t launch_head_bwd(const uint16_t* x, int ldx, const float* w, int C, long P, const float* dout, float* dx,
                           int lddx, float* dw, float* db, float* part, hipStream_t st) {
  // dw and db are adjacent in the parameter vector (db = dw + C)
  if (db != dw + C) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_head_bwd, dim3(kHeadGradParts), dim3(256), 0, st, x, ldx, w, C, P, dout, dx, lddx, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum_parts(part, kHeadGradParts, C + 1, dw, st);
}
hipError_t launch_flip_weights(const uint16_t* w, uint16_t* wt, int taps, int Cout, int Cin, hipStream_t st) {
  hipLaunchKernelGGL(k_flip_weights, dim3(blocks_for((long)taps * Cout * Cin)), dim3(256), 0, st, w, wt, taps,
                     Cout, Cin);
  return hipGetLastError();
}
hipError_t launch_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, st, w, w16, g, m, v, n, lr, b1, b2, eps, c1, c2,
                     gscale);
  return hipGetLastError();
}
hipError_t launch_f32_to_bf16(const float* in, uint16_t* out, long n, int ldo, int co0, int C, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks_for(n)), dim3(256), 0, st, in, out, n, ldo, co0, C);
  return hipGetLastError();
}
hipError_t launch_mse(const float* pred, const float* target, long P, float* dpred, double* loss, float scale,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_mse, dim3(blocks_for(P, 1024)), dim3(256), 0, st, pred, target, P, dpred, loss, scale);
  return hipGetLastError();
}
hipError_t launch_image_rms(const float* in, long P, int B, double* part, float* scale, hipStream_t st) {
  hipLaunchKernelGGL(k_image_rms_part, dim3(kRmsParts, B), dim3(256), 0, st, in, P, part);
  hipLaunchKernelGGL(k_image_rms_final, dim3((B + 63) / 64), dim3(64), 0, st, part, P, B, scale);
  return hipGetLastError();
}
int image_rms_parts() { return kRmsParts; }
hipError_t launch_to_bf16_scaled(const float* in, const float* scale, long P, int B, uint16_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16_scaled, dim3(blocks_for(P * B)), dim3(256), 0, st, in, scale, P, P * B, out);
  return hipGetLastError();
}
hipError_t launch_residual_add(const float* m, const float* y, const float* scale, long P, int B, float gain,
                               float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_residual_add, dim3(blocks_for(P * B)), dim3(256), 0, st, m, y, scale, P, P * B, gain, out);
  return hipGetLastError();
}
hipError_t launch_zero32(float* p, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, dim3(blocks_for(n)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace unet
