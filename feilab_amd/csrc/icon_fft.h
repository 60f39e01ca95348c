// Device-side FFT building blocks for CDNA4 (gfx950).
//
// A line of N complex values (N a power of two, 128..2048) is transformed by a
// team of T = N/16 threads.  Each thread holds 16 values in registers.  The
// transform is a Stockham autosort sequence of radix-16 passes followed by one
// pass of radix N/16^k (2, 4, 8 or 16); each pass is an in-register DFT with
// twiddles from a per-size table (double-precision sin/cos rounded to float)
// and passes exchange data through LDS.  Unnormalised, sign DIR = -1 (forward,
// e^{-2 pi i}) or +1 (inverse).
//
// Register order conventions (t = thread index within the team, T = N/16):
//   input  : v[r]            holds element t + r*T                 (r < 16)
//   output : v[q*RL + r]     holds element t + q*T + r*(N/RL)      (RL = last radix)
// so both the first load and the last store of a team touch consecutive
// addresses across consecutive threads.
#pragma once
#include <hip/hip_runtime.h>

namespace icon {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos / sin of 2 pi q / 16, q = 0..7.
constexpr float kCos16[8] = {1.0f, 0.923879532511286756f, 0.707106781186547524f,
                                        0.382683432365089772f, 0.0f, -0.382683432365089772f,
                                        -0.707106781186547524f, -0.923879532511286756f};
constexpr float kSin16[8] = {0.0f, 0.382683432365089772f, 0.707106781186547524f,
                                        0.923879532511286756f, 1.0f, 0.923879532511286756f,
                                        0.707106781186547524f, 0.382683432365089772f};

// a * exp(DIR * 2 pi i * Q / 16), Q in [0, 8), with the trivial cases exact.
template <int Q, int DIR>
__device__ __forceinline__ float2 rot16(float2 a) {
  if constexpr (Q == 0) {
    return a;
  } else if constexpr (Q == 4) {
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  } else {
    constexpr float c = kCos16[Q];
    constexpr float s = DIR > 0 ? kSin16[Q] : -kSin16[Q];
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int LOG>
__host__ __device__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 0; b < LOG; b++) r |= ((i >> b) & 1) << (LOG - 1 - b);
  return r;
}

template <int R> struct Log2 { static constexpr int v = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4; };

// Stage of the radix-2 decimation-in-time network: butterflies of span LEN.
template <int R, int DIR, int LEN, int K>
__device__ __forceinline__ void dit_stage_k(float2* v) {
  if constexpr (K < LEN / 2) {
#pragma unroll
    for (int i = K; i < R; i += LEN) {
      float2 u = v[i];
      float2 t = rot16<(16 / LEN) * K, DIR>(v[i + LEN / 2]);
      v[i] = c_add(u, t);
      v[i + LEN / 2] = c_sub(u, t);
    }
    dit_stage_k<R, DIR, LEN, K + 1>(v);
  }
}

template <int R, int DIR, int LEN>
__device__ __forceinline__ void dit_stages(float2* v) {
  if constexpr (LEN <= R) {
    dit_stage_k<R, DIR, LEN, 0>(v);
    dit_stages<R, DIR, LEN * 2>(v);
  }
}

// In-register DFT of size R (2, 4, 8, 16), natural order in and out.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  constexpr int L = Log2<R>::v;
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int j = bitrev<L>(i);
    if (i < j) {
      float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
  dit_stages<R, DIR, 2>(v);
}

template <int N, int NS>
__host__ __device__ constexpr int pass_radix() {
  return (N / NS) >= 16 ? 16 : N / NS;
}

template <int N>
__host__ __device__ constexpr int last_radix() {
  int ns = 1;
  int r = 16;
  while (ns < N) {
    r = (N / ns) >= 16 ? 16 : N / ns;
    ns *= r;
  }
  return r;
}

// LDS layout of a line: element i at float2 offset i*ES, plus PN padding
// float2 after every 2^PSH elements (PN = 0: no padding).  The padding moves
// elements that a pass's stores or loads place 2^PSH apart onto different
// banks; its offset (i >> PSH) * PN splits into a per-thread base and a
// compile-time constant at every access of the passes below.
template <int ES, int PSH = 0, int PN = 0>
struct Lay {
  static constexpr int kES = ES;
  __host__ __device__ static constexpr int at(int i) { return i * ES + (PN ? (i >> PSH) * PN : 0); }
  // float2 needed for a line of n elements
  __host__ __device__ static constexpr int size(int n) { return at(n - 1) + 1; }
};

// Loads a pass's inputs from LDS (layout L).
template <int N, int RP, typename L>
__device__ __forceinline__ void pass_load(float2 (&v)[16], const float2* lds, int t) {
  constexpr int T = N / 16;
  constexpr int B = 16 / RP;
#pragma unroll
  for (int q = 0; q < B; q++)
#pragma unroll
    for (int r = 0; r < RP; r++) v[q * RP + r] = lds[L::at(t + q * T + r * (N / RP))];
}

// Twiddles of the passes after the first (at most two for N <= 4096), per
// pass the B groups x (RP - 1) factors this thread applies.  They depend only
// on the thread's index, so they are loaded up front and their latency
// overlaps the data loads and the first pass.
struct FftTwiddles {
  float2 w[2][16];
};

// Number of passes after the first.
template <int N>
__host__ __device__ constexpr int twiddled_passes() {
  int ns = 1, k = 0;
  while (ns < N) {
    if (ns > 1) k++;
    ns *= N / ns >= 16 ? 16 : N / ns;
  }
  return k;
}

// The twiddle table of size N holds, for direction d (0: DIR = -1, 1: DIR =
// +1), 32 twiddles per thread t < N/16 as 16 float4 pairs: pair i of thread
// t is float4 (d * 16 + i) * N/16 + t, and twiddle slot s = 2i + {0, 1} is
// P*16 + q*(RP-1) + r-1 = exp(DIR 2 pi i r k step / N) of twiddled pass P
// (sub-transform size NS, radix RP, k = (t + q N/16) mod NS, step =
// N/(NS RP)).  Consecutive threads read consecutive 16 bytes.
template <int N, int DIR>
__device__ __forceinline__ void load_twiddles(FftTwiddles& T, int t, const float2* __restrict__ tw) {
  constexpr int TH = N / 16;
  const float4* p = reinterpret_cast<const float4*>(tw) + (DIR > 0 ? 16 * TH : 0) + t;
#pragma unroll
  for (int P = 0; P < twiddled_passes<N>(); P++) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float4 v = p[(P * 8 + i) * TH];
      T.w[P][2 * i] = make_float2(v.x, v.y);
      T.w[P][2 * i + 1] = make_float2(v.z, v.w);
    }
  }
}

// Twiddle and DFT of a pass with radix RP whose sub-transforms have size NS
// (twiddle slot P of T when NS > 1).
template <int N, int RP, int NS, int DIR, int P>
__device__ __forceinline__ void pass_compute(float2 (&v)[16], const FftTwiddles& T) {
  constexpr int B = 16 / RP;
#pragma unroll
  for (int q = 0; q < B; q++) {
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < RP; r++) v[q * RP + r] = c_mul(v[q * RP + r], T.w[P][q * (RP - 1) + r - 1]);
    }
    dft<RP, DIR>(&v[q * RP]);
  }
}

// Stores a pass's outputs to LDS in Stockham order (layout L).
template <int N, int RP, int NS, typename L>
__device__ __forceinline__ void pass_store(const float2 (&v)[16], float2* lds, int t) {
  constexpr int T = N / 16;
  constexpr int B = 16 / RP;
#pragma unroll
  for (int q = 0; q < B; q++) {
    const int j = t + q * T;
    const int k = j & (NS - 1);
    const int base = (j - k) * RP + k;
#pragma unroll
    for (int r = 0; r < RP; r++) lds[L::at(base + r * NS)] = v[q * RP + r];
  }
}

// Barrier between the LDS exchanges of a line: the whole block, or only the
// wavefront when one wave owns the line (WAVE = 1).
template <int WAVE>
__device__ __forceinline__ void line_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Runs the passes from sub-transform size NS on; v holds that pass's inputs.
// Every thread of the block (WAVE = 0) or of the wave (WAVE = 1) must call it.
// DFT of size 16 (natural order in and out) of a line whose inputs 4..11 are
// zero: the even outputs are the 8-point DFT of (x0..x3, x12..x15), the odd
// ones the 8-point DFT of the same values times exp(DIR 2 pi i r / 16), r
// their index (r = 12..15 contributing -exp(DIR 2 pi i (r-8) / 16)).
template <int DIR>
__device__ __forceinline__ void dft16_half(float2 (&v)[16]) {
  float2 a[8], b[8];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    a[m] = v[m];
    a[m + 4] = v[m + 12];
  }
  b[0] = v[0];
  b[1] = rot16<1, DIR>(v[1]);
  b[2] = rot16<2, DIR>(v[2]);
  b[3] = rot16<3, DIR>(v[3]);
  b[4] = rot16<4, DIR>(v[12]);
  b[5] = rot16<5, DIR>(v[13]);
  b[6] = rot16<6, DIR>(v[14]);
  b[7] = rot16<7, DIR>(v[15]);
#pragma unroll
  for (int m = 4; m < 8; m++) b[m] = make_float2(-b[m].x, -b[m].y);
  dft<8, DIR>(a);
  dft<8, DIR>(b);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    v[2 * j] = a[j];
    v[2 * j + 1] = b[j];
  }
}

// HALF: the line's elements N/4 .. 3N/4-1 are zero (a zero-padded line
// centred on element 0), so the inputs 4..11 of the first pass are zero and
// are not read.
template <int N, typename L, int DIR, int NS, int WAVE = 0, int P = 0, bool HALF = false>
__device__ __forceinline__ void fft_from(float2 (&v)[16], float2* lds, int t, const FftTwiddles& T) {
  constexpr int RP = pass_radix<N, NS>();
  if constexpr (HALF && NS == 1 && RP == 16) dft16_half<DIR>(v);
  else pass_compute<N, RP, NS, DIR, P>(v, T);
  if constexpr (NS * RP < N) {
    pass_store<N, RP, NS, L>(v, lds, t);
    line_sync<WAVE>();
    constexpr int RN = pass_radix<N, NS * RP>();
    pass_load<N, RN, L>(v, lds, t);
    line_sync<WAVE>();
    fft_from<N, L, DIR, NS * RP, WAVE, (NS > 1 ? P + 1 : 0)>(v, lds, t, T);
  }
}

// Twiddle slot P (the pass after the first that uses it) of T only, as
// load_twiddles loads it.
template <int N, int DIR, int P>
__device__ __forceinline__ void load_twiddle_pass(FftTwiddles& T, int t, const float2* __restrict__ tw) {
  constexpr int TH = N / 16;
  if constexpr (P < twiddled_passes<N>()) {
    const float4* p = reinterpret_cast<const float4*>(tw) + (DIR > 0 ? 16 * TH : 0) + t;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float4 v = p[(P * 8 + i) * TH];
      T.w[P][2 * i] = make_float2(v.x, v.y);
      T.w[P][2 * i + 1] = make_float2(v.z, v.w);
    }
  }
}

// fft_from with each pass's twiddles loaded during the exchange before it,
// so that only one pass's twiddles are held at a time (fewer registers than
// loading them all up front).
template <int N, typename L, int DIR, int NS, int P = 0, bool HALF = false>
__device__ __forceinline__ void fft_from_lazy(float2 (&v)[16], float2* lds, int t, FftTwiddles& T,
                                              const float2* __restrict__ tw) {
  constexpr int RP = pass_radix<N, NS>();
  if constexpr (HALF && NS == 1 && RP == 16) dft16_half<DIR>(v);
  else pass_compute<N, RP, NS, DIR, P>(v, T);
  if constexpr (NS * RP < N) {
    constexpr int PN = NS > 1 ? P + 1 : 0;   // the next pass's twiddle slot
    load_twiddle_pass<N, DIR, PN>(T, t, tw);
    pass_store<N, RP, NS, L>(v, lds, t);
    line_sync<0>();
    constexpr int RN = pass_radix<N, NS * RP>();
    pass_load<N, RN, L>(v, lds, t);
    line_sync<0>();
    fft_from_lazy<N, L, DIR, NS * RP, PN>(v, lds, t, T, tw);
  }
}

// fft_regs of a zero-padded line: v[4..11] are zero and are not read.
template <int N, typename L, int DIR>
__device__ __forceinline__ void fft_regs_half(float2 (&v)[16], float2* lds, int t, const FftTwiddles& T) {
  static_assert(N >= 128, "the first pass is radix 16");
  fft_from<N, L, DIR, 1, 0, 0, true>(v, lds, t, T);
}

// Registers in (input order) -> registers out (output order).  lds: scratch
// for this line in layout L (L::size(N) float2); T from load_twiddles<N, DIR>.
template <int N, typename L, int DIR>
__device__ __forceinline__ void fft_regs(float2 (&v)[16], float2* lds, int t, const FftTwiddles& T) {
  fft_from<N, L, DIR, 1>(v, lds, t, T);
}

// The same, loading the twiddles itself.
template <int N, typename L, int DIR>
__device__ __forceinline__ void fft_regs(float2 (&v)[16], float2* lds, int t, const float2* __restrict__ tw) {
  FftTwiddles T;
  load_twiddles<N, DIR>(T, t, tw);
  fft_from<N, L, DIR, 1>(v, lds, t, T);
}

// Element index held in output slot i of thread t.
template <int N>
__device__ __forceinline__ int out_index(int t, int i) {
  constexpr int RL = last_radix<N>();
  constexpr int T = N / 16;
  return t + (i / RL) * T + (i % RL) * (N / RL);
}

// True if output slot i of every thread holds only elements in [N/4, 3N/4):
// the part of a line that cropping a zero-padded transform back to its
// centred N/2 or fewer elements never keeps (its outputs need not be formed).
template <int N>
__host__ __device__ constexpr bool out_slot_cropped(int i) {
  constexpr int RL = last_radix<N>();
  constexpr int T = N / 16;
  const int lo = (i / RL) * T + (i % RL) * (N / RL);
  return lo >= N / 4 && lo + T <= 3 * N / 4;
}

// Element index held in input slot r of thread t.
template <int N>
__device__ __forceinline__ int in_index(int t, int r) {
  return t + r * (N / 16);
}

// LDS line in natural order -> transformed line in natural order (in place,
// layout L).  Callers synchronise before (data present) and after (data
// consumed).
template <int N, typename L, int DIR>
__device__ __forceinline__ void fft_lds(float2* lds, int t, const float2* __restrict__ tw) {
  FftTwiddles T;
  load_twiddles<N, DIR>(T, t, tw);
  float2 v[16];
  pass_load<N, 16, L>(v, lds, t);
  __syncthreads();
  fft_regs<N, L, DIR>(v, lds, t, T);
  // the last pass_load is followed by a barrier inside fft_from, so the
  // line can be overwritten now
#pragma unroll
  for (int i = 0; i < 16; i++) lds[L::at(out_index<N>(t, i))] = v[i];
}

}  // namespace icon
