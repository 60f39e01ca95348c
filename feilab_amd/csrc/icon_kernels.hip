// CDNA4 (gfx950) kernels of the ICON reconstruction core.
//
// Data layout in HBM (S slices batched, complex = interleaved float2):
//   m      [S][N0][N1] f32       image, index X*N1 + Y (X = width, Y = depth)
//   rhat   [S][N0][N1] c64       gradient  A^H W (A m) - A^H W b
//   ahwb   [S][N0][N1] c64       A^H W b
//   tmp    [S][N0][n1] c64       the N0 non-zero grid rows after the y-pass
//   part   [S][M][3]   c64       interpolation partials, one per column tile
//   fw     [S][M]      c64       weighted node values (spread input)
//   b      [S][M]      c64       Fourier-slice data
// The oversampled grid (n0 x n1, x slow) never exists in HBM as a whole: the
// forward transform writes only the N0 rows that the zero padding leaves
// non-zero, and the column pass keeps its n0 x C tile in LDS where it is
// interpolated from (forward) or spread into (adjoint).
//
// Reference kernels replaced (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
//   k_row_fwd      nufft_fhatDivCk_assignGhat(_R2C) :754-809 + y half of cufftExecC2C
//                  (+ ICONUpdateCU_R / INFRUpdateCU_R :1597-1639 fused in front)
//   k_col_fwd      x half of cufftExecC2C + nufft_GPhi_p1CU_noPreCompute :669-700
//   k_combine      nufft_GPhi_p2CU :704-727, complexMulDoubleCU :1539-1549,
//                  alpha1/matrixAdd/alpha2 :1652-1703 (FP64), data residual
//   k_col_adj      cudaMemset + nufft_adjoint_fullpre_FPhi :1078-1102 + x half of
//                  the inverse cufftExecC2C
//   k_row_adj      y half of the inverse FFT + nufft_assignF_HAT :1104-1121 +
//                  complexMinusCU :1563-1573 + alpha0 :1641-1650 (FP64)
//   k_sino         nufft_fhatDivCk1d_assignGhat_1d :1432-1452 + 1D FFT +
//                  nufft_GPhi_1dCU :1454-1477
//   k_reproj       cunfft_adjoint_1d :1518-1536 (1D gather in ascending node
//                  order, inverse FFT, nufft_assignF_HAT_1d :1123-1135)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "icon_fft.h"
#include "icon_kernels.h"

namespace icon {

// ------------------------------------------------------------------ row pass
// One grid row of length NR per team of NR/16 threads; RPB rows per block.
// LDS rows carry one float2 of padding after every 32 elements: the first
// pass stores 16 consecutive elements per thread, which without it would put
// every second thread of a wavefront on the same banks.
using RowLay = Lay<1, 5, 1>;
template <int NR>
constexpr int kRowLds = RowLay::size(NR);


// Inputs 0..3 and 12..15 of thread t's first pass over image row X of slice
// s (the others are padding): CPLX, the complex image rhat; otherwise the
// real image m, with the ICON update m <- (float)(m - alpha Re(rhat)) and its
// threshold applied and stored first (UPD; Re(rhat) from a.grad).  All loads are issued before any
// value is used (clamped to column 0 where the input is padding, the value
// then dropped): a load under a per-lane condition would be compiled into a
// branch whose merge waits for it before the next load is issued.
template <int NR, bool CPLX, bool UPD>
__device__ __forceinline__ void row_fwd_inputs(float2 (&v)[16], const RowFwdArgs& a, int t, int s, int X,
                                               bool valid, int h1) {
  float cj[8], mj[8], gj[8];
  float2 rj[8];
  bool ok[8];
  int Yj[8];
  const size_t row0 = ((size_t)s * a.N0 + X) * a.N1;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int Y = (in_index<NR>(t, j < 4 ? j : j + 8) + h1) & (NR - 1);
    ok[j] = valid && Y < a.N1;
    Yj[j] = ok[j] ? Y : 0;
    cj[j] = a.ck[(size_t)X * a.N1 + Yj[j]];
    if constexpr (!CPLX) mj[j] = a.m[row0 + Yj[j]];
    if constexpr (CPLX) rj[j] = a.rhat[row0 + Yj[j]];
    if constexpr (UPD) gj[j] = a.grad[row0 + Yj[j]];
  }
  const double al = UPD ? a.alpha[s] : 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    asm volatile("" : "+v"(cj[j]));
    if constexpr (!CPLX) asm volatile("" : "+v"(mj[j]));
    if constexpr (CPLX) asm volatile("" : "+v"(rj[j].x), "+v"(rj[j].y));
    if constexpr (UPD) asm volatile("" : "+v"(gj[j]));
  }
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = make_float2(0.f, 0.f);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (!ok[j]) continue;
    const int r = j < 4 ? j : j + 8;
    if constexpr (!CPLX) {
      float mv = mj[j];
      if constexpr (UPD) {
        // m - al*Re(rhat) in double, stored as float, then the threshold
        float nv = (float)__fma_rn(-al, (double)gj[j], (double)mv);
        if (a.update == 2) {
          if (a.data_type == 1) { if (nv > a.thr) nv = a.thr; }
          else { if (nv < a.thr) nv = a.thr; }
        }
        a.m[row0 + Yj[j]] = nv;
        mv = nv;
      }
      v[r].x = mv / cj[j];
    } else {
      v[r].x = rj[j].x / cj[j];
      v[r].y = rj[j].y / cj[j];
    }
  }
}

template <int NR, int RPB>
__device__ __forceinline__ void row_fwd_body(const RowFwdArgs& a, int bid) {
  constexpr int T = NR / 16;
  __shared__ float2 lds[RPB * kRowLds<NR>];
  const int team = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int row = bid * RPB + team;
  const bool valid = row < a.rows;
  const int s = valid ? row / a.N0 : 0;
  const int X = valid ? row % a.N0 : 0;
  const int h1 = a.N1 >> 1;
  // the N1 data columns sit at y in [NR - N1/2, NR) and [0, N1 - N1/2), so
  // with N1 <= NR/2 the inputs 4..11 of every thread are padding
  float2 v[16];
  if (!a.cplx) {
    if (a.update) row_fwd_inputs<NR, false, true>(v, a, t, s, X, valid, h1);
    else row_fwd_inputs<NR, false, false>(v, a, t, s, X, valid, h1);
  } else {
    row_fwd_inputs<NR, true, false>(v, a, t, s, X, valid, h1);
  }
  FftTwiddles Tw;
  load_twiddles<NR, -1>(Tw, t, a.tw);
  fft_regs_half<NR, RowLay, -1>(v, lds + team * kRowLds<NR>, t, Tw);
  if (valid) {
    float2* o = a.out + (size_t)row * NR;
    const unsigned zw = (unsigned)(a.zhi - a.zlo);
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((unsigned)(out_index<NR>(t, i) - a.zlo) >= zw) o[out_index<NR>(t, i)] = v[i];
  }
}

template <int NR, int RPB>
__global__ __launch_bounds__(RPB * (NR / 16)) void k_row_fwd(RowFwdArgs a) {
  row_fwd_body<NR, RPB>(a, blockIdx.x);
}

// Packed launch over several series: block b runs block items[b].y of the
// series whose argument block is descs[items[b].x]; the per-launch fields
// (image type, update, threshold) come by value.
template <int NR, int RPB>
__global__ __launch_bounds__(RPB * (NR / 16)) void k_row_fwd_multi(const RowFwdArgs* __restrict__ descs,
                                                                  const int2* __restrict__ items, RowFwdLaunch L) {
  const int2 it = items[blockIdx.x];
  RowFwdArgs a = descs[it.x];
  a.cplx = L.cplx; a.update = L.update; a.data_type = L.data_type; a.thr = L.thr;
  row_fwd_body<NR, RPB>(a, it.y);
}


// FWD: also the first row pass of A rhat (what k_row_fwd does for a complex
// image): rhat / ck padded to NR and y-transformed, written to a.fwd_out.
// The inverse's output slot i and the forward's input slot r of a thread hold
// the same element t + k T, so the hand-over is a fixed register permutation.
// Output slots of the inverse row transform that the crop keeps: the image
// (MODE 0: A^H result), or rhat = image - ahwb with its |rhat|^2 added to acc
// and, FWD, rhat / ck placed in u as the forward transform's input.  The ck
// and ahwb values of all slots are loaded before any is used (clamped to
// column 0 where the slot is padding, the value then dropped): a load under
// a per-lane condition would be compiled into a branch whose merge waits for
// it before the next load is issued.
template <int NR, bool MODE1, bool FWD>
__device__ __forceinline__ void row_adj_outputs(const float2 (&v)[16], float2 (&u)[16], double& acc,
                                                const RowAdjArgs& a, int t, int s, int X, bool valid, int h1) {
  const size_t row0 = ((size_t)s * a.N0 + X) * a.N1;
  float cj[16];
  float2 wj[16];
  bool ok[16];
  int Yj[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (out_slot_cropped<NR>(i)) continue;   // padding columns, as N1 <= NR/2
    const int Y = (out_index<NR>(t, i) + h1) & (NR - 1);
    ok[i] = valid && Y < a.N1;
    Yj[i] = ok[i] ? Y : 0;
    cj[i] = a.ck[(size_t)X * a.N1 + Yj[i]];
    if constexpr (MODE1) wj[i] = a.ahwb[row0 + Yj[i]];
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (out_slot_cropped<NR>(i)) continue;
    asm volatile("" : "+v"(cj[i]));
    if constexpr (MODE1) asm volatile("" : "+v"(wj[i].x), "+v"(wj[i].y));
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (out_slot_cropped<NR>(i)) continue;
    if (!ok[i]) continue;
    const float c = cj[i];
    const float2 f = make_float2(v[i].x / c, v[i].y / c);
    if constexpr (!MODE1) {
      a.img[row0 + Yj[i]] = f;
    } else {
      const float2 r = make_float2(f.x - wj[i].x, f.y - wj[i].y);
      a.grad[row0 + Yj[i]] = r.x;   // the update reads only Re(rhat)
      acc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
      if constexpr (FWD) {
        constexpr int RL = last_radix<NR>();
        u[i / RL + (i % RL) * (16 / RL)] = make_float2(r.x / c, r.y / c);
      }
    }
  }
}

template <int NR, int RPB, bool FWD>
__device__ __forceinline__ void row_adj_body(const RowAdjArgs& a, int bid) {
  constexpr int T = NR / 16;
  __shared__ float2 lds[RPB * kRowLds<NR>];
  __shared__ double red[RPB * T];
  const int team = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int row = bid * RPB + team;
  const bool valid = row < a.rows;
  const int s = valid ? row / a.N0 : 0;
  const int X = valid ? row % a.N0 : 0;
  // twiddles loaded pass by pass during the FFTs' exchanges (one pass's at a
  // time in registers)
  FftTwiddles Tw;
  // the row (row 0 for a team past the last row, its results dropped): all
  // 16 loads unconditional, so that they are in flight together.  Buffer
  // loads over the row: a column of the zero band gets an offset past the
  // row, which the load returns as 0 without reading memory.
  float2 v[16];
  {
    const auto rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(a.in + (size_t)(valid ? row : 0) * NR),
                                                      (short)0, NR * (int)sizeof(float2), 0x00020000);
    const unsigned zw = (unsigned)(a.zhi - a.zlo);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int y = in_index<NR>(t, r);
      const unsigned off = (unsigned)(y - a.zlo) >= zw ? (unsigned)y * 8u : 0x80000000u;
      v[r] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    }
  }
  fft_from_lazy<NR, RowLay, 1, 1>(v, lds + team * kRowLds<NR>, t, Tw, a.tw);
  const int h1 = a.N1 >> 1;
  double acc = 0.0;
  float2 u[16];   // FWD: forward input, slot r holds element t + r T
#pragma unroll
  for (int r = 0; r < 16; r++) u[r] = make_float2(0.f, 0.f);
  if (a.mode == 0) row_adj_outputs<NR, false, false>(v, u, acc, a, t, s, X, valid, h1);
  else row_adj_outputs<NR, true, FWD>(v, u, acc, a, t, s, X, valid, h1);
  if (a.mode == 1) {
    red[threadIdx.x] = acc;
    __syncthreads();
    if (valid && t == 0) {
      double sum = 0.0;
      for (int i = 0; i < T; i++) sum += red[team * T + i];
      a.rowpart[row] = sum;
    }
  }
  if constexpr (FWD) {
    // the inverse's last LDS exchange ended with a barrier
    fft_from_lazy<NR, RowLay, -1, 1, 0, true>(u, lds + team * kRowLds<NR>, t, Tw, a.tw);
    if (valid) {
      float2* o = a.fwd_out + (size_t)row * NR;
      const unsigned zw = (unsigned)(a.zhi - a.zlo);
#pragma unroll
      for (int i = 0; i < 16; i++)
        if ((unsigned)(out_index<NR>(t, i) - a.zlo) >= zw) o[out_index<NR>(t, i)] = u[i];
    }
  }
}

// The launch bound asks for 2 waves per SIMD (up to 256 registers); with the
// twiddles loaded pass by pass the kernel needs 112 and runs 4 waves per SIMD
// (with all twiddles up front it needed 132, and capped at 128 it spilled).
template <int NR, int RPB, bool FWD>
__global__ __launch_bounds__(RPB * (NR / 16), 2) void k_row_adj(RowAdjArgs a) {
  row_adj_body<NR, RPB, FWD>(a, blockIdx.x);
}

template <int NR, int RPB, bool FWD>
__global__ __launch_bounds__(RPB * (NR / 16), 2) void k_row_adj_multi(const RowAdjArgs* __restrict__ descs,
                                                                     const int2* __restrict__ items, int mode) {
  const int2 it = items[blockIdx.x];
  RowAdjArgs a = descs[it.x];
  a.mode = mode;
  if (!FWD) a.fwd_out = nullptr;
  row_adj_body<NR, RPB, FWD>(a, it.y);
}

// --------------------------------------------------------------- column pass
// One block per (slice, tile of C grid columns); the n0 x C tile lives in LDS,
// row-major with a padded row stride of C+1 float2 so that rows map to
// different banks.  FFT: thread (c = tid % C, t = tid / C) works on column c.
// Interpolation: one bin entry (node, run of y-taps inside the tile) per
// thread, summing its 14 x C cells of the tile.
// Spread: see k_col_adj (row-owner gather, no atomics).

// A tile's bin records are interleaved by field so that a wavefront reading
// one field of consecutive entries reads contiguous memory: for the tile's
// entries [e0, e1), x-tap pair q of entry e is float2 number q*(e1-e0) + e-e0
// of the block at float e0*14, and y-tap quad c4 is float4 number
// c4*(e1-e0) + e-e0 of the block at float e0*C (K = window taps).
template <int K>
__device__ __forceinline__ const float2* tile_px(const ColArgs& a, int e0) {
  return reinterpret_cast<const float2*>(a.ent_px + (size_t)e0 * K);
}
template <int C>
__device__ __forceinline__ const float4* tile_py(const ColArgs& a, int e0) {
  return reinterpret_cast<const float4*>(a.ent_py + (size_t)e0 * C);
}

// Row stride (float2) of the LDS tile: C + 2 keeps rows 16-byte aligned for
// vector reads while two blocks still share a CU's LDS; otherwise C + 1.
template <int NC, int C>
__host__ __device__ constexpr int col_stride() {
  return NC * (C + 2) * 8 <= 81920 ? C + 2 : C + 1;
}

// Layouts of the tile during the x FFT's exchanges.  A wavefront holds 64/C
// threads per column; at C = 8 each of its 8-byte loads and stores touches 8
// rows of the C columns at once, and the layout puts those rows on disjoint
// banks (2 cycles per 64-lane access, the minimum):
// * interpolation: row stride C, 8 float2 of padding after every 16 rows;
// * spread: row stride C, 1 float2 of padding after every 4 rows, which also
//   keeps the spread's own stores conflict-free (each thread writes its two
//   whole rows, 2 apart across lanes, where the padding every 16 rows would
//   put 4 lanes on one bank).
// The stride C + 1 used before took 3 cycles for most FFT accesses.
template <int C>
using ColFftLayFwd = Lay<C, 4, 8>;
template <int C>
using ColFftLay = Lay<C, 2, 1>;

// float2 of LDS per column block: the tile, or an FFT layout if larger
template <int NC, int C>
__host__ __device__ constexpr int col_lds() {
  constexpr int tile = NC * col_stride<NC, C>();
  constexpr int fwd = ColFftLayFwd<C>::at(NC - 1) + C, adj = ColFftLay<C>::at(NC - 1) + C;
  return tile > fwd ? (tile > adj ? tile : adj) : (fwd > adj ? fwd : adj);
}

template <int NC, int C, int K>
__device__ __forceinline__ void col_fwd_body(const ColArgs& a, int bid) {
  constexpr int NT = C * (NC / 16);
  constexpr int LS = col_stride<NC, C>();
  extern __shared__ __align__(16) float2 tile[];   // col_lds<NC, C>()
  // tile-major order: blocks in flight together work on the same column
  // tile of different slices and share its bin records in L2 (an order that
  // put the two tiles sharing each 128-byte grid line on one XCD back to back
  // halved the L2 misses but ran 3 % slower: the repeated lines come from
  // the memory-side cache).  (The spread's order by entry count measured no
  // faster here: an interpolation block's time varies far less with its
  // tile's entries.)
  const int blk = bid / a.S;
  const int s = bid % a.S;
  const int y0 = blk * C;
  const int tid = threadIdx.x;
  const int h0 = a.N0 >> 1;
  // a tile no node reaches (beyond the largest |ky| of the tilt range) has
  // nothing to interpolate: its FFT is not needed
  if (a.ent_off[blk] == a.ent_off[blk + 1]) return;
  // x FFT of column c = tid % C: the first pass takes its 16 inputs (grid
  // rows t + r NC/16; zero padding outside the N0 non-zero rows) straight
  // from global memory, all loads in flight together; the transform leaves
  // its output in the LDS tile for the interpolation
  {
    const int c = tid % C, t = tid / C;
    float2 v[16];
    // rows x in [NC/4, 3NC/4) (inputs 4..11) are padding, as N0 <= NC/2;
    // the other inputs are loaded unconditionally (rows past N0, only when
    // N0 < NC/2, read row 0 and are dropped afterwards).  A conditional load
    // per input would be compiled into branches whose merges wait for each
    // load before the next is issued; the empty asm below takes the loaded
    // values unconditionally, so that all loads are issued first.
    float2 g[8];
    bool ok[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int X = (in_index<NC>(t, j < 4 ? j : j + 8) + h0) & (NC - 1);
      ok[j] = X < a.N0;
      g[j] = a.in[((size_t)s * a.N0 + (ok[j] ? X : 0)) * a.n1 + y0 + c];
    }
    // the twiddles (first used after the first pass) are requested after
    // the grid rows, so that the first pass waits for the rows only
    FftTwiddles T;
    load_twiddles<NC, -1>(T, t, a.tw);
#pragma unroll
    for (int j = 0; j < 8; j++) asm volatile("" : "+v"(g[j].x), "+v"(g[j].y));
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (ok[j]) v[j < 4 ? j : j + 8] = g[j];
    fft_regs_half<NC, ColFftLayFwd<C>, -1>(v, tile + c, t, T);
    // the last exchange ended with a barrier: the column can be overwritten
#pragma unroll
    for (int i = 0; i < 16; i++) tile[out_index<NC>(t, i) * LS + c] = v[i];
  }
  __syncthreads();
  // interpolation partials of this tile's nodes: one bin entry per thread;
  // the entry's record (first x-tap, 14 x-taps, C y-taps) comes in with
  // vector loads and the thread sums its 14 x C cells from the LDS tile
  const int e0 = a.ent_off[blk], e1 = a.ent_off[blk + 1], ne = e1 - e0;
  for (int e = e0 + tid; e < e1; e += NT) {
    const int4 en = a.fent[e];
    float px[K];
    const float2* pxv = reinterpret_cast<const float2*>(a.fent_px + (size_t)e0 * K) + (e - e0);
#pragma unroll
    for (int q = 0; q < K / 2; q++) {
      const float2 p = pxv[q * ne];
      px[2 * q] = p.x;
      px[2 * q + 1] = p.y;
    }
    float py[C];
    const float4* pyv = reinterpret_cast<const float4*>(a.fent_py + (size_t)e0 * C) + (e - e0);
#pragma unroll
    for (int c4 = 0; c4 < C / 4; c4++) {
      const float4 v = pyv[c4 * ne];
      py[4 * c4] = v.x; py[4 * c4 + 1] = v.y; py[4 * c4 + 2] = v.z; py[4 * c4 + 3] = v.w;
    }
    // separable window: each grid row is first reduced over its C y-taps,
    // then weighted by its x-tap
    float ax = 0.f, ay = 0.f;
    auto add_row = [&](const float2* row, int q) {
      float2 gv[C];
      if constexpr (LS % 2 == 0) {
#pragma unroll
        for (int c2 = 0; c2 < C / 2; c2++) {
          const float4 g4 = reinterpret_cast<const float4*>(row)[c2];
          gv[2 * c2] = make_float2(g4.x, g4.y);
          gv[2 * c2 + 1] = make_float2(g4.z, g4.w);
        }
      } else {
#pragma unroll
        for (int c = 0; c < C; c++) gv[c] = row[c];
      }
      float tx = 0.f, ty = 0.f;
#pragma unroll
      for (int c = 0; c < C; c++) {
        tx = __fmaf_rn(gv[c].x, py[c], tx);
        ty = __fmaf_rn(gv[c].y, py[c], ty);
      }
      ax = __fmaf_rn(tx, px[q], ax);
      ay = __fmaf_rn(ty, px[q], ay);
    };
    // Entries whose taps wrap past row NC-1 have the largest first taps and
    // come last in the tile's order: only the last wavefronts can hold them;
    // the others take the fully unrolled loop with one base address and
    // constant row offsets.
    if (__builtin_amdgcn_ballot_w64(en.z > NC - K) == 0) {
      const float2* base = tile + en.z * LS;
#pragma unroll
      for (int q = 0; q < K; q++) add_row(base + q * LS, q);
    } else {
#pragma unroll 2
      for (int q = 0; q < K; q++) add_row(tile + ((en.z + q) & (NC - 1)) * LS, q);
    }
    a.part[((size_t)s * a.M + en.x) * kMaxSlotK + ((en.y >> 16) & 0xff)] = make_float2(ax, ay);
  }
}

template <int NC, int C, int K>
__global__ __launch_bounds__(C * (NC / 16), 4) void k_col_fwd(ColArgs a) {
  col_fwd_body<NC, C, K>(a, blockIdx.x);
}

template <int NC, int C, int K>
__global__ __launch_bounds__(C * (NC / 16), 4) void k_col_fwd_multi(const ColArgs* __restrict__ descs,
                                                                   const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const ColArgs a = descs[it.x];
  col_fwd_body<NC, C, K>(a, it.y);
}

template <int NC, int C, int K>
__device__ __forceinline__ void col_adj_body(const ColArgs& a, int bid) {
  constexpr int NT = C * (NC / 16);
  constexpr int RL = NC / NT;                    // adjacent grid rows per thread
  static_assert(RL >= 1 && RL <= 2 && NT * RL == NC, "one or two grid rows per thread");
  // staged record, 16-byte aligned fields: the K x-taps at [RL-1, RL-1+K)
  // with zeros around them in [0, XS), the y-taps at [XS, XS+C), then
  // (f.x, f.y, first x-tap, pad)
  constexpr int XS = (K + RL - 1 + 3) / 4 * 4;
  constexpr int REC = XS + C + 4;
  constexpr int CAP = col_lds<NC, C>() * 2 / REC;  // records staged at once in the tile's LDS
  extern __shared__ __align__(16) float2 tile[];   // col_lds<NC, C>()
  // tile-major order: blocks in flight together work on the same column
  // tile of different slices and share its bin records in L2; the tiles by
  // entry count, largest first, so that the dense tiles' long blocks start
  // first and short ones fill the end of the launch
  const int blk = a.tile_order[bid / a.S];
  const int s = bid % a.S;
  const int y0 = blk * C;
  const int tid = threadIdx.x;
  // a tile no node reaches: its spread grid and so its transform are zero;
  // inside the zero band the row kernel takes the zeros without loading them
  if (a.ent_off[blk] == a.ent_off[blk + 1]) {
    if (y0 >= a.zlo && y0 < a.zhi) return;
    for (int i = tid; i < a.N0 * C; i += NT)
      a.out[((size_t)s * a.N0 + i / C) * a.n1 + y0 + i % C] = make_float2(0.f, 0.f);
    return;
  }
  // Spread as a gather.  Rows [x, x+RL) of the tile are covered by the
  // entries with first x-tap z in [x-13, x+RL-1] (mod n0): one or two index
  // ranges of the tile's z-sorted list (run 0: the entries whose taps wrap
  // from the top rows, for the first K-1 rows only, else empty; run 1: the
  // rest).  The block stages the tile's records in LDS (CAP at a time) and
  // thread tid adds the entries of its rows x0 = RL tid, two per step, each
  // record's taps shared by its RL rows.  No atomics; every cell sums its
  // entries in list order.
  //
  // Dense tiles (at least a.split_min entries: the tiles around ky = 0,
  // where every tilt line reaches the rows around kx = 0 and their threads
  // hold up to 13x the tile's mean entries) split each row pair's entries,
  // run 0 then run 1, in two halves: thread tid adds the first half of its
  // own rows' into acc and the second half of its partner's (thread
  // tid ^ NT/2, rows x0 ^ NC/2, light where tid's are heavy) into accp.
  // A cell there is (its first half in list order) + (its second half in
  // list order).
  const int* xo = a.xoff + (size_t)blk * (NC + 1);
  const float2* fw = a.fw + (size_t)s * a.E;
  float* stage = reinterpret_cast<float*>(tile);
  const int x0 = tid * RL, xp = x0 ^ (NC / 2);
  const int e0 = a.ent_off[blk], e1 = a.ent_off[blk + 1];
  // the runs of the rows at x; the loads unconditional, so that they are in
  // flight together
  auto runs = [&](int x, int (&l)[2], int (&h)[2]) {
    const bool wrap = x < K - 1;
    l[0] = xo[wrap ? NC - (K - 1) + x : NC];
    h[0] = xo[NC];
    l[1] = xo[wrap ? 0 : x - (K - 1)];
    h[1] = xo[x + RL];
  };
  // adds entry record `rec` to the rows at x (accumulators ac)
  auto add = [&](const float* rec, int x, float2 (&ac)[RL][C]) {
    const float4 tail = reinterpret_cast<const float4*>(rec)[XS / 4 + C / 4];
    // x-tap slot of row x: (x - z) + RL - 1, in [0, K + RL - 2] by the range
    const int q = (x - __float_as_int(tail.z) + RL - 1) & (NC - 1);
    float px[RL];
#pragma unroll
    for (int r = 0; r < RL; r++) px[r] = rec[q + r];
    float py[C];
#pragma unroll
    for (int c4 = 0; c4 < C / 4; c4++) {
      const float4 v = reinterpret_cast<const float4*>(rec)[XS / 4 + c4];
      py[4 * c4] = v.x; py[4 * c4 + 1] = v.y; py[4 * c4 + 2] = v.z; py[4 * c4 + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < RL; r++) {
      // the value weighted by this row's x-tap, then by each y-tap
      const float fx = tail.x * px[r], fy = tail.y * px[r];
#pragma unroll
      for (int c = 0; c < C; c++) {
        ac[r][c].x = __fmaf_rn(fx, py[c], ac[r][c].x);
        ac[r][c].y = __fmaf_rn(fy, py[c], ac[r][c].y);
      }
    }
  };
  // adds the staged entries [lo, hi) of the chunk at cb (cn records) to the rows at x
  auto add_range = [&](int lo, int hi, int cb, int cn, int x, float2 (&ac)[RL][C]) {
    const int k1 = min(hi, cb + cn) - cb;
    int k = max(lo, cb) - cb;
    for (; k + 1 < k1; k += 2) {
      add(stage + k * REC, x, ac);
      add(stage + (k + 1) * REC, x, ac);
    }
    if (k < k1) add(stage + k * REC, x, ac);
  };
  FftTwiddles Tw;   // of the inverse x FFT, loaded while the tile is written
  auto gather = [&](auto split_tag) {
    constexpr bool SPLIT = decltype(split_tag)::value;
    int lo[2], hi[2], lp[2], hp[2];
    runs(x0, lo, hi);
    if constexpr (SPLIT) {
      runs(xp, lp, hp);
      // own rows: the first ceil(n/2) entries; the partner's: the rest of its
      const int n0 = hi[0] - lo[0], h = (n0 + hi[1] - lo[1] + 1) >> 1;
      hi[0] = lo[0] + min(n0, h);
      hi[1] = lo[1] + max(0, h - n0);
      const int m0 = hp[0] - lp[0], g = (m0 + hp[1] - lp[1] + 1) >> 1;
      lp[0] += min(m0, g);
      lp[1] += max(0, g - m0);
    }
    float2 acc[RL][C], accp[SPLIT ? RL : 1][C];
#pragma unroll
    for (int r = 0; r < RL; r++) {
#pragma unroll
      for (int c = 0; c < C; c++) acc[r][c] = make_float2(0.f, 0.f);
    }
    if constexpr (SPLIT) {
#pragma unroll
      for (int r = 0; r < RL; r++) {
#pragma unroll
        for (int c = 0; c < C; c++) accp[r][c] = make_float2(0.f, 0.f);
      }
    }
    for (int cb = e0; cb < e1; cb += CAP) {
      const int cn = min(CAP, e1 - cb);
      if (cb > e0) __syncthreads();   // the previous chunk has been consumed
      for (int i = tid; i < cn; i += NT) {
        const int e = cb + i;
        float4* rec = reinterpret_cast<float4*>(stage + i * REC);
        const float2* pxv = tile_px<K>(a, e0) + (e - e0);
        float t[XS];
#pragma unroll
        for (int q = 0; q < XS; q++) t[q] = 0.f;
#pragma unroll
        for (int q = 0; q < K / 2; q++) {
          const float2 v = pxv[q * (e1 - e0)];
          t[RL - 1 + 2 * q] = v.x;
          t[RL + 2 * q] = v.y;
        }
#pragma unroll
        for (int q = 0; q < XS / 4; q++) rec[q] = make_float4(t[4 * q], t[4 * q + 1], t[4 * q + 2], t[4 * q + 3]);
        const float4* pyv = tile_py<C>(a, e0) + (e - e0);
#pragma unroll
        for (int c4 = 0; c4 < C / 4; c4++) rec[XS / 4 + c4] = pyv[c4 * (e1 - e0)];
        const float2 f = fw[e];
        rec[XS / 4 + C / 4] = make_float4(f.x, f.y, __int_as_float(a.ent[e].z), 0.f);
      }
      __syncthreads();
#pragma unroll
      for (int run = 0; run < 2; run++) {
        add_range(lo[run], hi[run], cb, cn, x0, acc);
        if constexpr (SPLIT) add_range(lp[run], hp[run], cb, cn, xp, accp);
      }
    }
    if constexpr (SPLIT) {
      __syncthreads();   // the staged records are consumed
      // the partner's second halves into the tile, then each thread adds
      // its own first halves to them
#pragma unroll
      for (int r = 0; r < RL; r++) {
#pragma unroll
        for (int c = 0; c < C; c++) tile[ColFftLay<C>::at(xp + r) + c] = accp[r][c];
      }
      load_twiddles<NC, 1>(Tw, tid / C, a.tw);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RL; r++) {
#pragma unroll
        for (int c = 0; c < C; c++) {
          float2& cell = tile[ColFftLay<C>::at(x0 + r) + c];
          const float2 b = cell;
          cell = make_float2(acc[r][c].x + b.x, acc[r][c].y + b.y);
        }
      }
    } else {
      load_twiddles<NC, 1>(Tw, tid / C, a.tw);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RL; r++) {
#pragma unroll
        for (int c = 0; c < C; c++) tile[ColFftLay<C>::at(x0 + r) + c] = acc[r][c];
      }
    }
  };
  if (e1 - e0 >= a.split_min) gather(std::true_type{});
  else gather(std::false_type{});
  __syncthreads();
  // inverse x FFT of column c = tid % C; the last pass writes the N0 rows the
  // crop keeps (grid row x -> image row X = x + N0/2 mod NC) straight to global
  {
    const int c = tid % C, t = tid / C;
    float2 v[16];
    pass_load<NC, 16, ColFftLay<C>>(v, tile + c, t);
    __syncthreads();
    fft_regs<NC, ColFftLay<C>, 1>(v, tile + c, t, Tw);
    const int h0 = a.N0 >> 1;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (out_slot_cropped<NC>(i)) continue;   // padding rows, as N0 <= NC/2
      const int X = (out_index<NC>(t, i) + h0) & (NC - 1);
      if (X < a.N0) a.out[((size_t)s * a.N0 + X) * a.n1 + y0 + c] = v[i];
    }
  }
}

template <int NC, int C, int K>
__global__ __launch_bounds__(C * (NC / 16), 4) void k_col_adj(ColArgs a) {
  col_adj_body<NC, C, K>(a, blockIdx.x);
}

template <int NC, int C, int K>
__global__ __launch_bounds__(C * (NC / 16), 4) void k_col_adj_multi(const ColArgs* __restrict__ descs,
                                                                   const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const ColArgs a = descs[it.x];
  col_adj_body<NC, C, K>(a, it.y);
}

// ------------------------------------------------------------ node combine
// One block (1024 threads) per slice.  FP64 block reductions in a fixed order.


__device__ __forceinline__ double block_sum_1024(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int h = 512; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void combine_body(const CombineArgs& a, int s) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double acc = 0.0;
  const float2* P = a.part + (size_t)s * a.M * kMaxSlotK;
  for (int j = tid; j < a.M; j += 1024) {
    float2 f;
    if (a.mode == 2) {
      f = a.b[(size_t)s * a.M + j];
    } else {
      const int ns = a.nslot[j];
      f = P[(size_t)j * kMaxSlotK];
      for (int q = 1; q < ns; q++) {
        const float2 p = P[(size_t)j * kMaxSlotK + q];
        f.x += p.x;
        f.y += p.y;
      }
    }
    const float wj = a.w[j];
    if (a.mode == 0) {
      if (a.resid) {
        const float2 bj = a.b[(size_t)s * a.M + j];
        const double dx = (double)f.x - (double)bj.x, dy = (double)f.y - (double)bj.y;
        acc += (double)wj * (dx * dx + dy * dy);
      }
      if (a.frep) {
        const int i = j / a.N0;
        if (i == a.iomit) a.frep[(size_t)s * a.N0 + (j - i * a.N0)] = f;
      }
      const float2 v = make_float2(f.x * wj, f.y * wj);
      for (int q = 0; q < kMaxSlotK; q++) {
        const int e = a.ent_of_node[(size_t)j * kMaxSlotK + q];
        if (e >= 0) a.fw[(size_t)s * a.E + e] = v;
      }
    } else if (a.mode == 1) {
      acc += (double)f.x * (double)f.x * (double)wj + (double)f.y * (double)f.y * (double)wj;
    } else if (a.mode == 2) {
      acc += (double)wj * ((double)f.x * (double)f.x + (double)f.y * (double)f.y);
      const float2 v = make_float2(f.x * wj, f.y * wj);
      for (int q = 0; q < kMaxSlotK; q++) {
        const int e = a.ent_of_node[(size_t)j * kMaxSlotK + q];
        if (e >= 0) a.fw[(size_t)s * a.E + e] = v;
      }
    } else {
      a.fw[(size_t)s * a.M + j] = f;
    }
  }
  if (a.mode == 3) return;
  if (a.mode == 0 && !a.resid) return;
  const double tot = block_sum_1024(acc, red);
  if (a.mode == 0) {
    if (tid == 0) {
      const double bn = a.bnorm[s];
      a.resid[(size_t)s * a.resid_stride + a.resid_index] = bn > 0 ? sqrt(tot / bn) : 0.0;
    }
  } else if (a.mode == 2) {
    if (tid == 0) a.bnorm[s] = tot;
  } else {
    // rtr = sum of the row partials in row order, summed from LDS (red is
    // free after block_sum_1024's last barrier): the block loads them in one
    // round trip, where a loop of global loads on one thread waits for each
    double rtr = 0.0;
    for (int X0 = 0; X0 < a.N0; X0 += 1024) {
      if (X0 + tid < a.N0) red[tid] = a.rowpart[(size_t)s * a.N0 + X0 + tid];
      __syncthreads();
      if (tid == 0)
        for (int X = 0; X < min(1024, a.N0 - X0); X++) rtr += red[X];
      __syncthreads();
    }
    if (tid == 0) {
      const double al = tot != 0.0 ? rtr / tot : 0.0;
      a.alpha[s] = al;
      if (a.alpha_trace) a.alpha_trace[(size_t)s * a.alpha_stride + a.alpha_index] = al;
    }
  }
}

__global__ __launch_bounds__(1024) void k_combine(CombineArgs a) { combine_body(a, blockIdx.x); }

// Packed launch: the per-launch mode and trace indices by value; with_resid /
// with_frep switch the series' residual trace and reprojection values on.
__global__ __launch_bounds__(1024) void k_combine_multi(const CombineArgs* __restrict__ descs,
                                                        const int2* __restrict__ items, CombineLaunch L) {
  const int2 it = items[blockIdx.x];
  CombineArgs a = descs[it.x];
  a.mode = L.mode;
  a.resid_index = L.resid_index;
  a.alpha_index = L.alpha_index;
  if (!L.with_resid) a.resid = nullptr;
  if (!L.with_frep) a.frep = nullptr;
  combine_body(a, it.y);
}

// ------------------------------------------------------- projection rows (1D)
// One team of NR/16 threads per (slice, tilt) row; RPB rows per block.


template <int NR, int RPB>
__device__ __forceinline__ void sino_body(const SinoArgs& a, int bid) {
  constexpr int T = NR / 16;
  __shared__ float2 lds[RPB * kRowLds<NR>];
  const int team = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int row = bid * RPB + team;
  const bool valid = row < a.rows;
  const int h = a.N0 >> 1;
  // node row i of slice s reads input row i, or (cross validation) the rows
  // around the left-out one, the last node row being the data-free
  // reprojection line
  int in_row = row;
  bool has_data = valid;
  if (a.skip >= 0) {
    const int s = row / a.nz, i = row % a.nz;
    has_data = valid && i < a.nz - 1;
    in_row = s * a.nz + (i < a.skip ? i : i + 1);
  }
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    v[r] = make_float2(0.f, 0.f);
    if (r >= 4 && r < 12) continue;   // padding, as N0 <= NR/2
    const int x = in_index<NR>(t, r);
    const int X = (x + h) & (NR - 1);
    float gv = 0.f;
    if (has_data && X < a.N0) gv = a.sino[(size_t)in_row * a.N0 + X] / a.ck1[X];
    v[r] = make_float2(gv, 0.f);
  }
  float2* L = lds + team * kRowLds<NR>;
  FftTwiddles Tw;
  load_twiddles<NR, -1>(Tw, t, a.tw);
  fft_regs_half<NR, RowLay, -1>(v, L, t, Tw);
#pragma unroll
  for (int i = 0; i < 16; i++) L[out_index<NR>(t, i)] = v[i];
  __syncthreads();
  if (valid) {
    for (int k = t; k < a.N0; k += T) {
      float ax = 0.f, ay = 0.f;
      const int i0 = a.i10[k];
#pragma unroll
      for (int q = 0; q < a.K; q++) {
        const float ph = a.phi1[(size_t)k * a.K + q];
        const float2 gv = L[(i0 + q) & (NR - 1)];
        ax = __fmaf_rn(gv.x, ph, ax);
        ay = __fmaf_rn(gv.y, ph, ay);
      }
      a.b[(size_t)row * a.N0 + k] = make_float2(ax, ay);
    }
  }
}

template <int NR, int RPB>
__global__ __launch_bounds__(RPB * (NR / 16)) void k_sino(SinoArgs a) {
  sino_body<NR, RPB>(a, blockIdx.x);
}

template <int NR, int RPB>
__global__ __launch_bounds__(RPB * (NR / 16)) void k_sino_multi(const SinoArgs* __restrict__ descs,
                                                                const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const SinoArgs a = descs[it.x];
  sino_body<NR, RPB>(a, it.y);
}

// --------------------------------------------------------- reprojection (1D)
// One team of NR/16 threads per slice.  The 1D spread is a gather per grid
// cell over the nodes that touch it, in ascending node order.


template <int NR>
__device__ __forceinline__ void reproj_body(const ReprojArgs& a, int s) {
  __shared__ float2 lds[kRowLds<NR>];
  const int t = threadIdx.x;
  const float2* f = a.frep + (size_t)s * a.N0;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int x = in_index<NR>(t, r);
    float ax = 0.f, ay = 0.f;
    // Nodes touching cell x sit within a few cells of x (unwrapped as x - NR
    // for nodes near k = 0, as x otherwise); node spacing is sigma >= 2 cells,
    // so +-10 nodes around each candidate covers every tap.  The two windows
    // are visited in ascending k without overlap; membership is the exact
    // wrapped-tap test.
    int kprev = -1;
    for (int pass = 0; pass < 2; pass++) {
      const int u = pass == 0 ? x - NR : x;
      const int kc = (int)floorf((float)u / a.sigma) + (a.N0 >> 1);
      const int klo = max(max(kc - 10, 0), kprev + 1), khi = min(kc + 10, a.N0 - 1);
      for (int k = klo; k <= khi; k++) {
        const int d = (x - a.i10[k]) & (NR - 1);
        if (d < a.K) {
          const float ph = a.phi1[(size_t)k * a.K + d];
          ax = __fmaf_rn(f[k].x, ph, ax);
          ay = __fmaf_rn(f[k].y, ph, ay);
        }
      }
      kprev = max(kprev, khi);
    }
    v[r] = make_float2(ax, ay);
  }
  fft_regs<NR, RowLay, 1>(v, lds, t, a.tw);
  const int h = a.N0 >> 1;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (out_slot_cropped<NR>(i)) continue;   // padding, as N0 <= NR/2
    const int y = out_index<NR>(t, i);
    const int X = (y + h) & (NR - 1);
    if (X < a.N0) a.out[(size_t)s * a.N0 + X] = v[i].x / a.ck1[X];
  }
}

template <int NR>
__global__ __launch_bounds__(NR / 16) void k_reproj(ReprojArgs a) {
  reproj_body<NR>(a, blockIdx.x);
}

template <int NR>
__global__ __launch_bounds__(NR / 16) void k_reproj_multi(const ReprojArgs* __restrict__ descs,
                                                         const int2* __restrict__ items) {
  const int2 it = items[blockIdx.x];
  const ReprojArgs a = descs[it.x];
  reproj_body<NR>(a, it.y);
}

// -------------------------------------------------------------- plain FFTs

template <int NR, int RPB>
__global__ __launch_bounds__(RPB * (NR / 16)) void k_fft_rows(float2* data, int rows, int sign,
                                                             const float2* tw) {
  constexpr int T = NR / 16;
  __shared__ float2 lds[RPB * kRowLds<NR>];
  const int team = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int row = blockIdx.x * RPB + team;
  const bool valid = row < rows;
  float2* p = data + (size_t)(valid ? row : 0) * NR;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = valid ? p[in_index<NR>(t, r)] : make_float2(0.f, 0.f);
  if (sign < 0) fft_regs<NR, RowLay, -1>(v, lds + team * kRowLds<NR>, t, tw);
  else fft_regs<NR, RowLay, 1>(v, lds + team * kRowLds<NR>, t, tw);
  if (valid) {
#pragma unroll
    for (int i = 0; i < 16; i++) p[out_index<NR>(t, i)] = v[i];
  }
}

template <int NC, int C>
__global__ __launch_bounds__(C * (NC / 16)) void k_fft_cols(float2* data, int n1, int count, int sign,
                                                            const float2* tw) {
  constexpr int NT = C * (NC / 16);
  extern __shared__ float2 tile[];
  const int nblk = n1 / C;
  const int g = blockIdx.x / nblk;
  const int y0 = (blockIdx.x % nblk) * C;
  float2* base = data + (size_t)g * NC * n1;
  const int tid = threadIdx.x;
  constexpr int LS = C + 1;
  for (int i = tid; i < NC * C; i += NT) tile[(i / C) * LS + i % C] = base[(size_t)(i / C) * n1 + y0 + i % C];
  __syncthreads();
  if (sign < 0) fft_lds<NC, Lay<LS>, -1>(tile + (tid % C), tid / C, tw);
  else fft_lds<NC, Lay<LS>, 1>(tile + (tid % C), tid / C, tw);
  __syncthreads();
  for (int i = tid; i < NC * C; i += NT) base[(size_t)(i / C) * n1 + y0 + i % C] = tile[(i / C) * LS + i % C];
  (void)count;
}

__global__ void k_node_to_ent(const float2* __restrict__ f, float2* __restrict__ fw, const int4* __restrict__ ent,
                              int S, int M, int E) {
  const size_t n = (size_t)S * E;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(i / E), e = (int)(i % E);
    fw[i] = f[(size_t)s * M + ent[e].x];
  }
}

__global__ void k_zero_f32(float* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

__global__ void k_copy_f32(float* dst, const float* src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Plain 16-byte vector copy (bandwidth / counter calibration).
__global__ __launch_bounds__(256) void k_copy_f4(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ================================================================ launchers

namespace {

thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;

template <int NR>
constexpr int rows_per_block() {
  return NR >= 2048 ? 2 : 256 / (NR / 16) > 16 ? 16 : 256 / (NR / 16);
}

// The adjoint row kernel (two transforms per row, a block barrier between
// every exchange) runs in blocks of 128 threads: fewer wavefronts wait at
// each barrier (0.833 vs 0.874 ms at 256 x 512 rows of 1024 against 256
// threads; 64 threads: 0.840 ms).
template <int NR>
constexpr int row_adj_rows_per_block() {
  return 128 / (NR / 16) > 1 ? 128 / (NR / 16) : 1;
}

}  // namespace

void set_launch_events(hipEvent_t start, hipEvent_t stop) {
  t_ev_start = start;
  t_ev_stop = stop;
}

// hipLaunchKernelGGL, or, while timing events are set, hipExtLaunchKernelGGL
// with them (the events then bracket the kernel itself, not the enqueue)
#define LAUNCH_TIMED(KERNEL, GRID, BLOCK, SHM, ST, ...)                                          \
  do {                                                                                           \
    if (t_ev_start)                                                                              \
      hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, ST, t_ev_start, t_ev_stop, 0, __VA_ARGS__); \
    else                                                                                         \
      hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, ST, __VA_ARGS__);                             \
  } while (0)

#define ROW_CASES(X) X(128) X(256) X(512) X(1024) X(2048)

hipError_t launch_row_fwd(int NR, const RowFwdArgs& a, hipStream_t st) {
  const RowFwdArgs& h = a;
  switch (NR) {
#define CASE(N)                                                                                     \
  case N: {                                                                                         \
    constexpr int RPB = rows_per_block<N>();                                                        \
    LAUNCH_TIMED((k_row_fwd<N, RPB>), dim3((h.rows + RPB - 1) / RPB), dim3(RPB * (N / 16)), 0, \
                       st, a);                                                                      \
    break;                                                                                          \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_row_adj(int NR, const RowAdjArgs& a, hipStream_t st) {
  const RowAdjArgs& h = a;
  switch (NR) {
#define CASE(N)                                                                                     \
  case N: {                                                                                         \
    constexpr int RPB = row_adj_rows_per_block<N>();                                                \
    if (h.fwd_out)                                                                                  \
      LAUNCH_TIMED((k_row_adj<N, RPB, true>), dim3((h.rows + RPB - 1) / RPB), dim3(RPB * (N / 16)), 0, \
                         st, a);                                                                    \
    else                                                                                            \
      LAUNCH_TIMED((k_row_adj<N, RPB, false>), dim3((h.rows + RPB - 1) / RPB), dim3(RPB * (N / 16)), 0, \
                         st, a);                                                                    \
    break;                                                                                          \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sino(int NR, const SinoArgs& a, hipStream_t st) {
  const SinoArgs& h = a;
  switch (NR) {
#define CASE(N)                                                                                  \
  case N: {                                                                                      \
    constexpr int RPB = rows_per_block<N>();                                                     \
    LAUNCH_TIMED((k_sino<N, RPB>), dim3((h.rows + RPB - 1) / RPB), dim3(RPB * (N / 16)), 0, \
                       st, a);                                                                   \
    break;                                                                                       \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reproj(int NR, const ReprojArgs& a, hipStream_t st) {
  const ReprojArgs& h = a;
  switch (NR) {
#define CASE(N)                                                                   \
  case N:                                                                         \
    hipLaunchKernelGGL((k_reproj<N>), dim3(h.S), dim3(N / 16), 0, st, a);         \
    break;
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fft_rows(int NR, float2* data, int rows, int sign, const float2* tw, hipStream_t st) {
  switch (NR) {
#define CASE(N)                                                                                     \
  case N: {                                                                                         \
    constexpr int RPB = rows_per_block<N>();                                                        \
    hipLaunchKernelGGL((k_fft_rows<N, RPB>), dim3((rows + RPB - 1) / RPB), dim3(RPB * (N / 16)), 0, \
                       st, data, rows, sign, tw);                                                   \
    break;                                                                                          \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Column kernels: (NC, C) with NC * C * 8 bytes of LDS.
// (n0, C, taps) of the column kernels built: the Gaussian window at every
// tile width, the Kaiser-Bessel window at C = 8
#define COL_CASES(X)                                                                                 \
  X(128, 8, 14) X(128, 16, 14) X(256, 8, 14) X(256, 16, 14) X(512, 8, 14) X(512, 16, 14) X(1024, 8, 14) \
  X(1024, 16, 14) X(2048, 8, 14) X(128, 8, 6) X(256, 8, 6) X(512, 8, 6) X(1024, 8, 6) X(2048, 8, 6)
#define FFT_COL_CASES(X) X(128, 8) X(128, 16) X(256, 8) X(256, 16) X(512, 8) X(512, 16) X(1024, 8) X(1024, 16) X(2048, 8)

bool col_tile_ok(int NC, int n1, int C, int K) {
  // the interpolation's block order pairs tiles in groups of 16 blocks
  return col_config_ok(NC, C, K) && n1 % C == 0 && (n1 / C) % 16 == 0;
}

bool col_config_ok(int NC, int C, int K) {
#define CASE(N, CC, KK) if (NC == N && C == CC && K == KK) return true;
  COL_CASES(CASE)
#undef CASE
  return false;
}

template <int NC, int C, int K>
static hipError_t col_attr() {
  static bool done = false;
  if (!done) {
    hipError_t e = hipFuncSetAttribute((const void*)k_col_fwd<NC, C, K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       col_lds<NC, C>() * (int)sizeof(float2));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_col_adj<NC, C, K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            col_lds<NC, C>() * (int)sizeof(float2));
    if (e != hipSuccess) return e;
    done = true;
  }
  return hipSuccess;
}

template <int NC, int C>
static hipError_t fft_col_attr() {
  static bool done = false;
  if (!done) {
    hipError_t e = hipFuncSetAttribute((const void*)k_fft_cols<NC, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       NC * (C + 1) * (int)sizeof(float2));
    if (e != hipSuccess) return e;
    done = true;
  }
  return hipSuccess;
}

hipError_t set_col_attrs(int NC, int C, int K) {
#define CASE(N, CC, KK) \
  if (NC == N && C == CC && K == KK) return col_attr<N, CC, KK>();
  COL_CASES(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

hipError_t launch_col(int NC, int C, bool adjoint, const ColArgs& a, hipStream_t st) {
  const ColArgs& h = a;
  const int nblk = h.n1 / C;
  if (!col_tile_ok(NC, h.n1, C, h.K)) return hipErrorInvalidValue;
  hipError_t e = hipErrorInvalidValue;
#define CASE(N, CC, KK)                                                                                  \
  if (NC == N && C == CC && h.K == KK) {                                                                 \
    e = col_attr<N, CC, KK>();                                                                           \
    if (e != hipSuccess) return e;                                                                       \
    if (adjoint)                                                                                         \
      LAUNCH_TIMED((k_col_adj<N, CC, KK>), dim3(h.S * nblk), dim3(CC * (N / 16)),                  \
                         (col_lds<N, CC>()) * sizeof(float2), st, a);                                    \
    else                                                                                                 \
      LAUNCH_TIMED((k_col_fwd<N, CC, KK>), dim3(h.S * nblk), dim3(CC * (N / 16)),                  \
                         (col_lds<N, CC>()) * sizeof(float2), st, a);                                    \
    return hipGetLastError();                                                                            \
  }
  COL_CASES(CASE)
#undef CASE
  return e;
}

hipError_t launch_fft_cols(int NC, int C, float2* data, int n1, int count, int sign, const float2* tw,
                           hipStream_t st) {
  const int nblk = n1 / C;
#define CASE(N, CC)                                                                                     \
  if (NC == N && C == CC) {                                                                             \
    hipError_t e = fft_col_attr<N, CC>();                                                                   \
    if (e != hipSuccess) return e;                                                                      \
    hipLaunchKernelGGL((k_fft_cols<N, CC>), dim3(count * nblk), dim3(CC * (N / 16)), N * (CC + 1) * sizeof(float2), \
                       st, data, n1, count, sign, tw);                                                  \
    return hipGetLastError();                                                                           \
  }
  FFT_COL_CASES(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------- packed launches

int row_fwd_rows_per_block(int NR) {
  switch (NR) {
#define CASE(N) case N: return rows_per_block<N>();
    ROW_CASES(CASE)
#undef CASE
    default: return 0;
  }
}

int row_adj_rows_per_block(int NR) {
  switch (NR) {
#define CASE(N) case N: return row_adj_rows_per_block<N>();
    ROW_CASES(CASE)
#undef CASE
    default: return 0;
  }
}

int sino_rows_per_block(int NR) { return row_fwd_rows_per_block(NR); }

hipError_t launch_row_fwd_multi(int NR, const RowFwdArgs* descs, const int2* items, int n, const RowFwdLaunch& L,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (NR) {
#define CASE(N)                                                                                   \
  case N: {                                                                                       \
    constexpr int RPB = rows_per_block<N>();                                                      \
    LAUNCH_TIMED((k_row_fwd_multi<N, RPB>), dim3(n), dim3(RPB * (N / 16)), 0, st, descs, items, L); \
    break;                                                                                        \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_row_adj_multi(int NR, const RowAdjArgs* descs, const int2* items, int n, int mode, bool fwd,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (NR) {
#define CASE(N)                                                                                         \
  case N: {                                                                                             \
    constexpr int RPB = row_adj_rows_per_block<N>();                                                    \
    if (fwd)                                                                                            \
      LAUNCH_TIMED((k_row_adj_multi<N, RPB, true>), dim3(n), dim3(RPB * (N / 16)), 0, st, descs, items, mode); \
    else                                                                                                \
      LAUNCH_TIMED((k_row_adj_multi<N, RPB, false>), dim3(n), dim3(RPB * (N / 16)), 0, st, descs, items, mode); \
    break;                                                                                              \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sino_multi(int NR, const SinoArgs* descs, const int2* items, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (NR) {
#define CASE(N)                                                                                \
  case N: {                                                                                    \
    constexpr int RPB = rows_per_block<N>();                                                   \
    LAUNCH_TIMED((k_sino_multi<N, RPB>), dim3(n), dim3(RPB * (N / 16)), 0, st, descs, items);    \
    break;                                                                                     \
  }
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_reproj_multi(int NR, const ReprojArgs* descs, const int2* items, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (NR) {
#define CASE(N)                                                                          \
  case N:                                                                                \
    hipLaunchKernelGGL((k_reproj_multi<N>), dim3(n), dim3(N / 16), 0, st, descs, items); \
    break;
    ROW_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NC, int C, int K>
static hipError_t col_multi_attr() {
  static bool done = false;
  if (!done) {
    hipError_t e = hipFuncSetAttribute((const void*)k_col_fwd_multi<NC, C, K>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, col_lds<NC, C>() * (int)sizeof(float2));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)k_col_adj_multi<NC, C, K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            col_lds<NC, C>() * (int)sizeof(float2));
    if (e != hipSuccess) return e;
    done = true;
  }
  return hipSuccess;
}

hipError_t set_col_multi_attrs(int NC, int C, int K) {
#define CASE(N, CC, KK) \
  if (NC == N && C == CC && K == KK) return col_multi_attr<N, CC, KK>();
  COL_CASES(CASE)
#undef CASE
  return hipErrorInvalidValue;
}

hipError_t launch_col_multi(int NC, int C, int K, bool adjoint, const ColArgs* descs, const int2* items, int n,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipError_t e = hipErrorInvalidValue;
#define CASE(N, CC, KK)                                                                                   \
  if (NC == N && C == CC && K == KK) {                                                                    \
    e = col_multi_attr<N, CC, KK>();                                                                      \
    if (e != hipSuccess) return e;                                                                        \
    if (adjoint)                                                                                          \
      LAUNCH_TIMED((k_col_adj_multi<N, CC, KK>), dim3(n), dim3(CC * (N / 16)),                            \
                   (col_lds<N, CC>()) * sizeof(float2), st, descs, items);                                \
    else                                                                                                  \
      LAUNCH_TIMED((k_col_fwd_multi<N, CC, KK>), dim3(n), dim3(CC * (N / 16)),                            \
                   (col_lds<N, CC>()) * sizeof(float2), st, descs, items);                                \
    return hipGetLastError();                                                                             \
  }
  COL_CASES(CASE)
#undef CASE
  return e;
}

hipError_t launch_combine_multi(const CombineArgs* descs, const int2* items, int n, const CombineLaunch& L,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  LAUNCH_TIMED(k_combine_multi, dim3(n), dim3(1024), 0, st, descs, items, L);
  return hipGetLastError();
}

hipError_t launch_combine(const CombineArgs& a, hipStream_t st) {
  LAUNCH_TIMED(k_combine, dim3(a.S), dim3(1024), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_node_to_ent(const float2* f, float2* fw_ent, const int4* ent, int S, int M, int E, hipStream_t st) {
  const size_t n = (size_t)S * E;
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_node_to_ent, dim3(blocks), dim3(256), 0, st, f, fw_ent, ent, S, M, E);
  return hipGetLastError();
}

hipError_t launch_zero(float* p, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_f32, dim3(blocks), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_copy4(float4* dst, const float4* src, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 32);
  hipLaunchKernelGGL(k_copy_f4, dim3(blocks), dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_copy(float* dst, const float* src, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_copy_f32, dim3(blocks), dim3(256), 0, st, dst, src, n);
  return hipGetLastError();
}

}  // namespace icon
