// ICON reconstruction core: CPU restatement of the reference algorithm.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path
// is checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker / CPU baseline, never
// as the thing measured or shipped.
//
// Two precisions share one template:
//   prec 32  "FP32 mirror": float storage and arithmetic with the reference's
//            float/double placement (window tables and deapodisation in the
//            reference's mixed precision, FMA where the reference's compiler
//            contracts, FP64 step-size reductions, update in double stored as
//            float).
//   prec 64  "FP64 NUFFT oracle": the same algorithm with every table, FFT and
//            accumulation in double.
// Plus an exact O(M N0 N1) NDFT for tiny sizes.
//
// Reference files followed (paths relative to
// tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
//   MainCode.cu:38-109,113-160      setup and per-slice loop
//   Iteration2.cu:33-146            ICON/INFR iteration
//   alpha_new2.cu:28-114            exact line-search step
//   nufft_gpu_v8.cu:22-42           2D deapodisation ck
//   nufft_gpu_v8.cu:44-163          adjoint CSR (cell -> ascending node order)
//   nufft_gpu_v8.cu:221-235         Gaussian window taps and wrapped indices
//   nufft_gpu_v8.cu:445-556         2D plan
//   nufft_gpu_v8.cu:669-727         interpolation (two phases)
//   nufft_gpu_v8.cu:754-809         deapodise + zero-pad
//   nufft_gpu_v8.cu:984-1076        forward NUFFT of a real image
//   nufft_gpu_v8.cu:1078-1121       adjoint gather + crop/deapodise
//   nufft_gpu_v8.cu:1167-1238       adjoint NUFFT
//   nufft_gpu_v8.cu:1242-1536       1D plan / forward / adjoint
//   nufft_gpu_v8.cu:1539-1673       ICON pointwise kernels and reductions
//   px2D.cu:8-47, px2D.cuh:8        radial node coordinates
//   weight.cu:8-41                  density weights
//   readSlicePart.cu:10-22          Fourier slice data b
//   reProject_NFFT.cu:81-98         reprojection at the min-|theta| tilt
//
// Parity status of each function is stated in its comment; the pinning tests
// live in tests/test_oracle.py.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <complex>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPI = 3.1415926535898;      // nufft_gpu_v8.cuh:8
constexpr double kPiPx = 3.14159265;         // px2D.cuh:8
constexpr int kTaps = 14;                    // 2m+2 with m = NFFTM = 6 (nufft_gpu_v8.cuh:5)

// ---------------------------------------------------------------- geometry

// Oversampling factor and grid size, MainCode.cu:66-71.  The reference keeps
// uptop and thita in float and takes n = ceil(thita*N) in float arithmetic.
void grid_size(int N, float* sigma, int* n) {
  float uptop = (float)std::pow(2.0, std::ceil(std::log10((double)N) / std::log10(2.0)));
  float thita = (float)(2.0 * uptop / (double)N);
  float prod = thita * (float)N;
  *sigma = thita;
  *n = (int)std::ceil(prod);
}

// Window shape parameters, nufft_gpu_v8.cu:471-472 (float b, bf).
void window_params_f32(float thita, int m, float* b, float* bf) {
  float num = 2.0f * thita * (float)m;
  double den = (double)(2.0f * thita - 1.0f) * kPI;
  *b = (float)((double)num / den);
  *bf = (float)(1.0 / std::sqrt(kPI * (double)*b));
}

// Density weights, weight.cu:8-41.  theta are the (negated) tilt angles in
// degrees; w has nz rows of nx.
void weight(float* w, int nx, int nz, const float* th) {
  int center = nx / 2;
  for (int i = 0; i < nz; i++) {
    for (int j = 0; j < nx; j++) {
      float v;
      if (j - center == 0) {
        v = (float)(180.0 / (4 * nz));
      } else if (i == 0) {
        v = (float)(2.0 * std::fabs((double)((th[i + 1] - th[i]) * (float)(j - center))));
      } else if (i == nz - 1) {
        v = (float)(2.0 * std::fabs((double)((th[i - 1] - th[i]) * (float)(j - center))));
      } else {
        v = (float)(1.0 * std::fabs((double)((th[i - 1] - th[i + 1]) * (float)(j - center))));
      }
      w[i * nx + j] = v;
    }
  }
  // A single tilt leaves only the branches above that index i+1 / i-1; the
  // reference reads past the array there.  nz == 1 is rejected by the callers.
}

// Radial node coordinates, px2D.cu:8-47 (TFunction2): f accumulates in float
// from -0.5 in steps of (float)(1/N); x = f cos(theta pi/180), y = f sin(...)
// in double, stored as float.  Node j = i*nx + k (tilt-major).
void px2d(float* xy, int nx, int nz, const float* th) {
  std::vector<float> f(nx);
  float x1 = -0.5f;
  float step = (float)(1.0 / nx);
  f[0] = x1;
  for (int i = 1; i < nx; i++) f[i] = f[i - 1] + step;
  for (int t = 0; t < nz; t++) {
    double a = (double)th[t] * kPiPx / 180.0;
    double c = std::cos(a), s = std::sin(a);
    for (int i = 0; i < nx; i++) {
      xy[2 * (t * nx + i)] = (float)((double)f[i] * c);
      xy[2 * (t * nx + i) + 1] = (float)((double)f[i] * s);
    }
  }
}

// 1D nodes of the projection transform, MainCode.cu:75-78.
void px1d(float* x, int nx) {
  for (int k = 0; k < nx; k++)
    x[k] = (float)(((double)(float)k - (double)(float)nx / 2.0) / (double)(float)nx);
}

// ---------------------------------------------------------------- FFT

// Unnormalised radix-2 DFT, sign -1 (forward, e^{-2 pi i}) or +1 (inverse).
// Twiddles from double-precision sin/cos rounded to T.  n must be a power of 2.
template <class T>
void fft1(std::complex<T>* a, int n, int sign) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    for (int k = 0; k < half; k++) {
      double ang = sign * 2.0 * M_PI * k / len;
      std::complex<T> w((T)std::cos(ang), (T)std::sin(ang));
      for (int i = k; i < n; i += len) {
        std::complex<T> u = a[i];
        std::complex<T> v = a[i + half];
        T vr = v.real() * w.real() - v.imag() * w.imag();
        T vi = v.real() * w.imag() + v.imag() * w.real();
        a[i] = std::complex<T>(u.real() + vr, u.imag() + vi);
        a[i + half] = std::complex<T>(u.real() - vr, u.imag() - vi);
      }
    }
  }
}

// 2D transform of an n0 x n1 row-major grid (index x*n1 + y): y first, then x.
template <class T>
void fft2(std::complex<T>* g, int n0, int n1, int sign) {
  for (int x = 0; x < n0; x++) fft1(g + (size_t)x * n1, n1, sign);
  std::vector<std::complex<T>> col(n0);
  for (int y = 0; y < n1; y++) {
    for (int x = 0; x < n0; x++) col[x] = g[(size_t)x * n1 + y];
    fft1(col.data(), n0, sign);
    for (int x = 0; x < n0; x++) g[(size_t)x * n1 + y] = col[x];
  }
}

template <class T> inline T fmaT(T a, T b, T c);
template <> inline float fmaT<float>(float a, float b, float c) { return std::fmaf(a, b, c); }
template <> inline double fmaT<double>(double a, double b, double c) { return std::fma(a, b, c); }

// ---------------------------------------------------------------- plan

template <class T>
struct Plan {
  int N0, N1, nz, M, n0, n1, m;
  float sig0, sig1;
  T b0, bf0, b1, bf1;
  std::vector<float> theta;       // negated angles
  std::vector<float> nodes;       // 2*M (x, y)
  std::vector<float> w;           // M
  std::vector<T> ck;              // N0*N1, index X*N1+Y
  std::vector<T> phix, phiy;      // M*14
  std::vector<int> ix, iy;        // M*14 wrapped grid indices
  // 1D projection plan (N0 nodes, grid n0)
  std::vector<T> ck1;             // N0
  std::vector<T> phi1;            // N0*14
  std::vector<int> i1;            // N0*14
  std::vector<float> px1;
  int iomit;                      // index of the first min-|theta| tilt
  float omit;
  int prec;
};

// Window taps for one coordinate, nufft_gpu_v8.cu:221-235 (float mirror).
void phi_taps_f32(const float* coord, int stride, int M, int N, float thita, float b,
                  float bf, int m, float* phi, int* idx) {
  int n = (int)std::ceil((float)N * thita);
  for (int j = 0; j < M; j++) {
    float c = coord[(size_t)j * stride];
    for (int t = 0; t < 2 * m + 2; t++) {
      int l = (int)(std::floor((float)n * c) - (float)m + (float)t);
      float x1 = c - (float)l / (float)n;
      float nx1 = (float)n * x1;
      phi[(size_t)j * (2 * m + 2) + t] = bf * std::exp(-(nx1 * nx1 / b));
      idx[(size_t)j * (2 * m + 2) + t] = (l + n) % n;
    }
  }
}

// FP64 window taps: same window, b and bf in double, sigma = n/N exactly.
void phi_taps_f64(const float* coord, int stride, int M, int n, double b, double bf, int m,
                  double* phi, int* idx) {
  for (int j = 0; j < M; j++) {
    double c = coord[(size_t)j * stride];
    for (int t = 0; t < 2 * m + 2; t++) {
      int l = (int)std::floor(n * c) - m + t;
      double x1 = c - (double)l / n;
      double nx1 = n * x1;
      phi[(size_t)j * (2 * m + 2) + t] = bf * std::exp(-(nx1 * nx1 / b));
      idx[(size_t)j * (2 * m + 2) + t] = ((l % n) + n) % n;
    }
  }
}

// 1D deapodisation, nufft_gpu_v8.cu:1258-1271 / one factor of :22-42: computed
// in double with the float b, stored in float.
float ck_factor_f32(int k, int N, float thita, int m) {
  float nn = std::ceil((float)N * thita);
  float b = (float)((double)(2.0f * thita * (float)m) / ((double)(2.0f * thita - 1.0f) * kPI));
  double a = kPI * (double)k / (double)nn;
  return (float)std::exp(-(double)b * (a * a));
}

double ck_factor_f64(int k, int n, double b) {
  double a = M_PI * (double)k / (double)n;
  return std::exp(-b * a * a);
}

template <class T>
Plan<T>* plan_create(int N0, int N1, int nz, const float* theta, int m, int prec) {
  auto* p = new Plan<T>();
  p->N0 = N0; p->N1 = N1; p->nz = nz; p->M = N0 * nz; p->m = m; p->prec = prec;
  grid_size(N0, &p->sig0, &p->n0);
  grid_size(N1, &p->sig1, &p->n1);
  p->theta.assign(theta, theta + nz);
  // omit angle: MainCode.cu:46-49 (strictly smaller |theta| replaces)
  float om = theta[0];
  for (int i = 0; i < nz; i++) if (std::fabs(om) > std::fabs(theta[i])) om = theta[i];
  p->omit = om;
  p->iomit = 0;
  for (int i = 0; i < nz; i++) if (theta[i] == om) { p->iomit = i; break; }
  const int M = p->M;
  p->nodes.resize(2 * (size_t)M);
  px2d(p->nodes.data(), N0, nz, theta);
  p->w.resize(M);
  weight(p->w.data(), N0, nz, theta);
  p->ck.resize((size_t)N0 * N1);
  p->phix.resize((size_t)M * kTaps); p->phiy.resize((size_t)M * kTaps);
  p->ix.resize((size_t)M * kTaps); p->iy.resize((size_t)M * kTaps);
  p->px1.resize(N0);
  px1d(p->px1.data(), N0);
  p->ck1.resize(N0);
  p->phi1.resize((size_t)N0 * kTaps); p->i1.resize((size_t)N0 * kTaps);
  if (prec == 32) {
    float b0, bf0, b1, bf1;
    window_params_f32(p->sig0, m, &b0, &bf0);
    window_params_f32(p->sig1, m, &b1, &bf1);
    p->b0 = b0; p->bf0 = bf0; p->b1 = b1; p->bf1 = bf1;
    std::vector<float> ph((size_t)M * kTaps);
    phi_taps_f32(p->nodes.data(), 2, M, N0, p->sig0, b0, bf0, m, ph.data(), p->ix.data());
    for (size_t i = 0; i < ph.size(); i++) p->phix[i] = ph[i];
    phi_taps_f32(p->nodes.data() + 1, 2, M, N1, p->sig1, b1, bf1, m, ph.data(), p->iy.data());
    for (size_t i = 0; i < ph.size(); i++) p->phiy[i] = ph[i];
    // ck2d2: nufft_gpu_v8.cu:22-42
    for (int X = 0; X < N0; X++) {
      float cx = ck_factor_f32(X - N0 / 2, N0, p->sig0, m);
      for (int Y = 0; Y < N1; Y++) {
        float cy = ck_factor_f32(Y - N1 / 2, N1, p->sig1, m);
        p->ck[(size_t)X * N1 + Y] = cx * cy;
      }
    }
    // 1D plan: nufft_gpu_v8.cu:1298-1391 (same sigma as dimension 0)
    std::vector<float> ph1((size_t)N0 * kTaps);
    phi_taps_f32(p->px1.data(), 1, N0, N0, p->sig0, b0, bf0, m, ph1.data(), p->i1.data());
    for (size_t i = 0; i < ph1.size(); i++) p->phi1[i] = ph1[i];
    for (int X = 0; X < N0; X++) p->ck1[X] = ck_factor_f32(X - N0 / 2, N0, p->sig0, m);
  } else {
    double s0 = (double)p->n0 / N0, s1 = (double)p->n1 / N1;
    double b0 = 2 * s0 * m / ((2 * s0 - 1) * M_PI), b1 = 2 * s1 * m / ((2 * s1 - 1) * M_PI);
    p->b0 = (T)b0; p->b1 = (T)b1;
    p->bf0 = (T)(1.0 / std::sqrt(M_PI * b0)); p->bf1 = (T)(1.0 / std::sqrt(M_PI * b1));
    std::vector<double> ph((size_t)M * kTaps);
    phi_taps_f64(p->nodes.data(), 2, M, p->n0, b0, p->bf0, m, ph.data(), p->ix.data());
    for (size_t i = 0; i < ph.size(); i++) p->phix[i] = (T)ph[i];
    phi_taps_f64(p->nodes.data() + 1, 2, M, p->n1, b1, p->bf1, m, ph.data(), p->iy.data());
    for (size_t i = 0; i < ph.size(); i++) p->phiy[i] = (T)ph[i];
    for (int X = 0; X < N0; X++)
      for (int Y = 0; Y < N1; Y++)
        p->ck[(size_t)X * N1 + Y] =
            (T)(ck_factor_f64(X - N0 / 2, p->n0, b0) * ck_factor_f64(Y - N1 / 2, p->n1, b1));
    std::vector<double> ph1((size_t)N0 * kTaps);
    phi_taps_f64(p->px1.data(), 1, N0, p->n0, b0, p->bf0, m, ph1.data(), p->i1.data());
    for (size_t i = 0; i < ph1.size(); i++) p->phi1[i] = (T)ph1[i];
    for (int X = 0; X < N0; X++) p->ck1[X] = (T)ck_factor_f64(X - N0 / 2, p->n0, b0);
  }
  return p;
}

using cd = std::complex<double>;

// Deapodise and zero-pad into the oversampled grid, nufft_gpu_v8.cu:754-809:
// image (X,Y) -> grid (x,y) with X = (x + N0/2) % n0 < N0.
template <class T>
void pad_deapod(const Plan<T>& p, const std::complex<T>* img, std::complex<T>* g) {
  const int N0 = p.N0, N1 = p.N1, n0 = p.n0, n1 = p.n1;
  for (int x = 0; x < n0; x++) {
    int X = (x + N0 / 2) % n0;
    for (int y = 0; y < n1; y++) {
      int Y = (y + N1 / 2) % n1;
      std::complex<T> v(0, 0);
      if (X < N0 && Y < N1) {
        T c = p.ck[(size_t)X * N1 + Y];
        const std::complex<T>& s = img[(size_t)X * N1 + Y];
        v = std::complex<T>(s.real() / c, s.imag() / c);
      }
      g[(size_t)x * n1 + y] = v;
    }
  }
}

// Two-phase interpolation, nufft_gpu_v8.cu:669-727: phase 1 sums the 14
// x-taps per (node, y-tap) with weight phix*phiy (FMA accumulate), phase 2
// sums the 14 phase-1 partials in order.
template <class T>
void interp(const Plan<T>& p, const std::complex<T>* g, std::complex<T>* f) {
  const int n1 = p.n1;
  for (int k = 0; k < p.M; k++) {
    T fx = 0, fy = 0;
    for (int ty = 0; ty < kTaps; ty++) {
      int iy = p.iy[(size_t)k * kTaps + ty];
      T py = p.phiy[(size_t)k * kTaps + ty];
      T ax = 0, ay = 0;
      for (int tx = 0; tx < kTaps; tx++) {
        int ix = p.ix[(size_t)k * kTaps + tx];
        T phi = p.phix[(size_t)k * kTaps + tx] * py;
        const std::complex<T>& gv = g[(size_t)ix * n1 + iy];
        ax = fmaT(gv.real(), phi, ax);
        ay = fmaT(gv.imag(), phi, ay);
      }
      fx += ax;
      fy += ay;
    }
    f[k] = std::complex<T>(fx, fy);
  }
}

// Forward NUFFT, nufft_gpu_v8.cu:984-1076 (real image) and :844-949 (complex).
template <class T>
void forward(const Plan<T>& p, const std::complex<T>* img, std::complex<T>* f,
             std::vector<std::complex<T>>& g) {
  g.resize((size_t)p.n0 * p.n1);
  pad_deapod(p, img, g.data());
  fft2(g.data(), p.n0, p.n1, -1);
  interp(p, g.data(), f);
}

// Adjoint gather, nufft_gpu_v8.cu:1078-1102 with the CSR of :44-163: every
// touched cell sums f[k]*(phix*phiy) over its nodes in ascending k (FMA).
// Looping k ascending and scattering gives the same per-cell order.
template <class T>
void spread(const Plan<T>& p, const std::complex<T>* f, std::complex<T>* g) {
  const int n1 = p.n1;
  std::fill(g, g + (size_t)p.n0 * p.n1, std::complex<T>(0, 0));
  for (int k = 0; k < p.M; k++) {
    const std::complex<T> fk = f[k];
    for (int ty = 0; ty < kTaps; ty++) {
      int iy = p.iy[(size_t)k * kTaps + ty];
      T py = p.phiy[(size_t)k * kTaps + ty];
      for (int tx = 0; tx < kTaps; tx++) {
        int ix = p.ix[(size_t)k * kTaps + tx];
        T phi = p.phix[(size_t)k * kTaps + tx] * py;
        std::complex<T>& c = g[(size_t)ix * n1 + iy];
        c = std::complex<T>(fmaT(fk.real(), phi, c.real()), fmaT(fk.imag(), phi, c.imag()));
      }
    }
  }
}

// Crop and deapodise, nufft_gpu_v8.cu:1104-1121.
template <class T>
void crop_deapod(const Plan<T>& p, const std::complex<T>* g, std::complex<T>* img) {
  const int N0 = p.N0, N1 = p.N1, n0 = p.n0, n1 = p.n1;
  for (int X = 0; X < N0; X++) {
    int x = (X + n0 - N0 / 2) % n0;
    for (int Y = 0; Y < N1; Y++) {
      int y = (Y + n1 - N1 / 2) % n1;
      T c = p.ck[(size_t)X * N1 + Y];
      const std::complex<T>& v = g[(size_t)x * n1 + y];
      img[(size_t)X * N1 + Y] = std::complex<T>(v.real() / c, v.imag() / c);
    }
  }
}

// Adjoint NUFFT, nufft_gpu_v8.cu:1167-1238.
template <class T>
void adjoint(const Plan<T>& p, const std::complex<T>* f, std::complex<T>* img,
             std::vector<std::complex<T>>& g) {
  g.resize((size_t)p.n0 * p.n1);
  spread(p, f, g.data());
  fft2(g.data(), p.n0, p.n1, +1);
  crop_deapod(p, g.data(), img);
}

// 1D forward NUFFT of one projection row, nufft_gpu_v8.cu:1432-1516.
template <class T>
void trafo1d(const Plan<T>& p, const std::complex<T>* in, std::complex<T>* out) {
  const int N = p.N0, n = p.n0;
  std::vector<std::complex<T>> g(n);
  for (int x = 0; x < n; x++) {
    int X = (x + N / 2) % n;
    if (X < N) {
      T c = p.ck1[X];
      g[x] = std::complex<T>(in[X].real() / c, in[X].imag() / c);
    } else {
      g[x] = std::complex<T>(0, 0);
    }
  }
  fft1(g.data(), n, -1);
  for (int k = 0; k < N; k++) {
    T fx = 0, fy = 0;
    for (int t = 0; t < kTaps; t++) {
      T ph = p.phi1[(size_t)k * kTaps + t];
      const std::complex<T>& gv = g[p.i1[(size_t)k * kTaps + t]];
      fx = fmaT(gv.real(), ph, fx);
      fy = fmaT(gv.imag(), ph, fy);
    }
    out[k] = std::complex<T>(fx, fy);
  }
}

// 1D adjoint NUFFT, nufft_gpu_v8.cu:1518-1536 with the 1D CSR of :58-99.
template <class T>
void adjoint1d(const Plan<T>& p, const std::complex<T>* f, std::complex<T>* out) {
  const int N = p.N0, n = p.n0;
  std::vector<std::complex<T>> g(n, std::complex<T>(0, 0));
  for (int k = 0; k < N; k++) {
    for (int t = 0; t < kTaps; t++) {
      T ph = p.phi1[(size_t)k * kTaps + t];
      std::complex<T>& c = g[p.i1[(size_t)k * kTaps + t]];
      c = std::complex<T>(fmaT(f[k].real(), ph, c.real()), fmaT(f[k].imag(), ph, c.imag()));
    }
  }
  fft1(g.data(), n, +1);
  for (int X = 0; X < N; X++) {
    int x = (X + n - N / 2) % n;
    T c = p.ck1[X];
    out[X] = std::complex<T>(g[x].real() / c, g[x].imag() / c);
  }
}

// Fourier-slice data of one slice: b_i = NFFT1D(p_i) for every tilt,
// readSlicePart.cu:10-22.  sino is nz rows of N0 (tilt-major).
template <class T>
void sino_to_b(const Plan<T>& p, const float* sino, std::complex<T>* b) {
  std::vector<std::complex<T>> row(p.N0);
  for (int i = 0; i < p.nz; i++) {
    for (int k = 0; k < p.N0; k++) row[k] = std::complex<T>((T)sino[(size_t)i * p.N0 + k], 0);
    trafo1d(p, row.data(), b + (size_t)i * p.N0);
  }
}

struct SolveOpts {
  int pre, icon, post, dataType;
  float thr;
};

// ICON solve of one slice, MainCode.cu:115-135 + Iteration2.cu:33-146 +
// alpha_new2.cu:28-114.  m_out is N0*N1 (index X*N1+Y).  resid (if given)
// receives R(m_i) for i = 0..iters (R = sqrt(sum w|Am-b|^2)/sqrt(sum w|b|^2),
// FP64), alpha (if given) the step of each iteration.
template <class T>
void icon_solve(const Plan<T>& p, const float* sino, const SolveOpts& o, float* m_out,
                double* resid, double* alpha_tr, const float* m_init) {
  const int M = p.M, NN = p.N0 * p.N1;
  using C = std::complex<T>;
  std::vector<C> b(M), f(M), ahwb(NN), rhat(NN), mimg(NN), g;
  sino_to_b(p, sino, b.data());
  double bnorm = 0;
  for (int j = 0; j < M; j++)
    bnorm += (double)p.w[j] * ((double)b[j].real() * b[j].real() + (double)b[j].imag() * b[j].imag());
  // wb, Ahwb (MainCode.cu:123-134)
  std::vector<C> wb(M);
  for (int j = 0; j < M; j++) wb[j] = C(b[j].real() * (T)p.w[j], b[j].imag() * (T)p.w[j]);
  adjoint(p, wb.data(), ahwb.data(), g);
  std::vector<T> m(NN, (T)0);   // Iteration2.cu:33
  if (m_init) for (int i = 0; i < NN; i++) m[i] = (T)m_init[i];
  const int iters = o.pre + o.icon + o.post;
  auto residual = [&](const std::vector<C>& fv) {
    double r = 0;
    for (int j = 0; j < M; j++) {
      double dx = (double)fv[j].real() - (double)b[j].real();
      double dy = (double)fv[j].imag() - (double)b[j].imag();
      r += (double)p.w[j] * (dx * dx + dy * dy);
    }
    return bnorm > 0 ? std::sqrt(r / bnorm) : 0.0;
  };
  for (int it = 0; it < iters; it++) {
    for (int i = 0; i < NN; i++) mimg[i] = C(m[i], 0);
    forward(p, mimg.data(), f.data(), g);                       // Iteration2.cu:53
    if (resid) resid[it] = residual(f);
    for (int j = 0; j < M; j++)                                 // :69
      f[j] = C(f[j].real() * (T)p.w[j], f[j].imag() * (T)p.w[j]);
    adjoint(p, f.data(), rhat.data(), g);                       // :80
    for (int i = 0; i < NN; i++)                                // :96
      rhat[i] = C(rhat[i].real() - ahwb[i].real(), rhat[i].imag() - ahwb[i].imag());
    // Alpha2: rtr, A rhat, rtar, al (alpha_new2.cu:28-114)
    double rtr = 0;
    for (int i = 0; i < NN; i++)
      rtr += (double)rhat[i].real() * (double)rhat[i].real() +
             (double)rhat[i].imag() * (double)rhat[i].imag();
    forward(p, rhat.data(), f.data(), g);
    double rtar = 0;
    for (int j = 0; j < M; j++)
      rtar += (double)f[j].real() * (double)f[j].real() * (double)p.w[j] +
              (double)f[j].imag() * (double)f[j].imag() * (double)p.w[j];
    double al = rtar != 0 ? rtr / rtar : 0.0;
    if (alpha_tr) alpha_tr[it] = al;
    // ICONUpdateCU_R / INFRUpdateCU_R, nufft_gpu_v8.cu:1597-1639
    bool clampit = (o.pre <= it && it < iters - o.post);
    for (int i = 0; i < NN; i++) {
      T v = (T)std::fma(-al, (double)rhat[i].real(), (double)m[i]);
      if (clampit) {
        if (o.dataType == 1) { if (v > (T)o.thr) v = (T)o.thr; }
        else { if (v < (T)o.thr) v = (T)o.thr; }
      }
      m[i] = v;
    }
  }
  if (resid) {
    for (int i = 0; i < NN; i++) mimg[i] = C(m[i], 0);
    forward(p, mimg.data(), f.data(), g);
    resid[iters] = residual(f);
  }
  for (int i = 0; i < NN; i++) m_out[i] = (float)m[i];
}

// Reprojection at the min-|theta| tilt, reProject_NFFT.cu:81-98: the forward
// transform on that tilt's nodes (identical to tilt iomit's nodes of the full
// plan), then the 1D adjoint, real part.
template <class T>
void reproject(const Plan<T>& p, const float* m, float* out) {
  const int NN = p.N0 * p.N1;
  using C = std::complex<T>;
  std::vector<C> img(NN), f(p.M), g, o(p.N0);
  for (int i = 0; i < NN; i++) img[i] = C((T)m[i], 0);
  forward(p, img.data(), f.data(), g);
  adjoint1d(p, f.data() + (size_t)p.iomit * p.N0, o.data());
  for (int X = 0; X < p.N0; X++) out[X] = (float)o[X].real();
}

struct AnyPlan {
  int prec;
  Plan<float>* p32;
  Plan<double>* p64;
};

template <class T> Plan<T>* get(AnyPlan* a);
template <> Plan<float>* get<float>(AnyPlan* a) { return a->p32; }
template <> Plan<double>* get<double>(AnyPlan* a) { return a->p64; }

template <class T>
void to_c(const double* in, std::complex<T>* out, size_t n) {
  for (size_t i = 0; i < n; i++) out[i] = std::complex<T>((T)in[2 * i], (T)in[2 * i + 1]);
}
template <class T>
void from_c(const std::complex<T>* in, double* out, size_t n) {
  for (size_t i = 0; i < n; i++) { out[2 * i] = in[i].real(); out[2 * i + 1] = in[i].imag(); }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int orc_grid(int N, float* sigma, int* n) { grid_size(N, sigma, n); return 0; }

void orc_window_params(float sigma, int m, float* b, float* bf) { window_params_f32(sigma, m, b, bf); }

void orc_weight(float* w, int nx, int nz, const float* theta) { weight(w, nx, nz, theta); }

void orc_px2d(float* xy, int nx, int nz, const float* theta) { px2d(xy, nx, nz, theta); }

void orc_px1d(float* x, int nx) { px1d(x, nx); }

// MainCode.cu:38-49: read whitespace-separated degrees, negate, find the
// min-|theta| angle.  Returns the count, or -1 if the file cannot be opened.
int orc_read_angles(const char* path, float* theta, int cap, float* omit) {
  FILE* fp = std::fopen(path, "r");
  if (!fp) return -1;
  int n = 0;
  float v;
  while (n < cap && std::fscanf(fp, "%f", &v) == 1) theta[n++] = -v;
  std::fclose(fp);
  if (n > 0) {
    float om = theta[0];
    for (int i = 0; i < n; i++) if (std::fabs(om) > std::fabs(theta[i])) om = theta[i];
    *omit = om;
  }
  return n;
}

void* orc_plan_create(int N0, int N1, int nz, const float* theta, int m, int prec) {
  auto* a = new AnyPlan();
  a->prec = prec;
  a->p32 = nullptr; a->p64 = nullptr;
  if (prec == 32) a->p32 = plan_create<float>(N0, N1, nz, theta, m, 32);
  else a->p64 = plan_create<double>(N0, N1, nz, theta, m, 64);
  return a;
}

void orc_plan_destroy(void* h) {
  auto* a = (AnyPlan*)h;
  delete a->p32; delete a->p64; delete a;
}

// out: N0, N1, nz, M, n0, n1, iomit
void orc_plan_info(void* h, int* out, float* sig_b_bf) {
  auto* a = (AnyPlan*)h;
  auto fill = [&](auto* p) {
    out[0] = p->N0; out[1] = p->N1; out[2] = p->nz; out[3] = p->M;
    out[4] = p->n0; out[5] = p->n1; out[6] = p->iomit;
    if (sig_b_bf) {
      sig_b_bf[0] = p->sig0; sig_b_bf[1] = (float)p->b0; sig_b_bf[2] = (float)p->bf0;
      sig_b_bf[3] = p->sig1; sig_b_bf[4] = (float)p->b1; sig_b_bf[5] = (float)p->bf1;
    }
  };
  if (a->p32) fill(a->p32); else fill(a->p64);
}

// Copies the plan's tables out (any pointer may be NULL).
void orc_plan_tables(void* h, double* ck, double* phix, double* phiy, int* ix, int* iy,
                     float* w, float* nodes, double* ck1, double* phi1, int* i1) {
  auto* a = (AnyPlan*)h;
  auto dump = [&](auto* p) {
    size_t MT = (size_t)p->M * kTaps;
    if (ck) for (size_t i = 0; i < p->ck.size(); i++) ck[i] = p->ck[i];
    if (phix) for (size_t i = 0; i < MT; i++) phix[i] = p->phix[i];
    if (phiy) for (size_t i = 0; i < MT; i++) phiy[i] = p->phiy[i];
    if (ix) std::memcpy(ix, p->ix.data(), MT * sizeof(int));
    if (iy) std::memcpy(iy, p->iy.data(), MT * sizeof(int));
    if (w) std::memcpy(w, p->w.data(), p->M * sizeof(float));
    if (nodes) std::memcpy(nodes, p->nodes.data(), 2 * (size_t)p->M * sizeof(float));
    if (ck1) for (int i = 0; i < p->N0; i++) ck1[i] = p->ck1[i];
    if (phi1) for (size_t i = 0; i < p->phi1.size(); i++) phi1[i] = p->phi1[i];
    if (i1) std::memcpy(i1, p->i1.data(), p->i1.size() * sizeof(int));
  };
  if (a->p32) dump(a->p32); else dump(a->p64);
}

// Complex buffers are interleaved (re, im) doubles.
void orc_forward(void* h, const double* img, double* f) {
  auto* a = (AnyPlan*)h;
  auto run = [&](auto* p) {
    using T = typename std::remove_pointer<decltype(p)>::type;
    using V = decltype(p->b0);
    (void)sizeof(T);
    size_t NN = (size_t)p->N0 * p->N1;
    std::vector<std::complex<V>> im(NN), fo(p->M), g;
    to_c(img, im.data(), NN);
    forward(*p, im.data(), fo.data(), g);
    from_c(fo.data(), f, p->M);
  };
  if (a->p32) run(a->p32); else run(a->p64);
}

void orc_adjoint(void* h, const double* f, double* img) {
  auto* a = (AnyPlan*)h;
  auto run = [&](auto* p) {
    using V = decltype(p->b0);
    size_t NN = (size_t)p->N0 * p->N1;
    std::vector<std::complex<V>> fi(p->M), im(NN), g;
    to_c(f, fi.data(), p->M);
    adjoint(*p, fi.data(), im.data(), g);
    from_c(im.data(), img, NN);
  };
  if (a->p32) run(a->p32); else run(a->p64);
}

// Fourier-slice data b (unweighted) from a sinogram of nz rows of N0 floats.
void orc_sino_to_b(void* h, const float* sino, double* b) {
  auto* a = (AnyPlan*)h;
  auto run = [&](auto* p) {
    using V = decltype(p->b0);
    std::vector<std::complex<V>> bo(p->M);
    sino_to_b(*p, sino, bo.data());
    from_c(bo.data(), b, p->M);
  };
  if (a->p32) run(a->p32); else run(a->p64);
}

void orc_adjoint1d(void* h, const double* f, double* out) {
  auto* a = (AnyPlan*)h;
  auto run = [&](auto* p) {
    using V = decltype(p->b0);
    std::vector<std::complex<V>> fi(p->N0), o(p->N0);
    to_c(f, fi.data(), p->N0);
    adjoint1d(*p, fi.data(), o.data());
    from_c(o.data(), out, p->N0);
  };
  if (a->p32) run(a->p32); else run(a->p64);
}

void orc_fft(int n, int sign, int prec, double* data) {
  if (prec == 32) {
    std::vector<std::complex<float>> v(n);
    to_c(data, v.data(), n);
    fft1(v.data(), n, sign);
    from_c(v.data(), data, n);
  } else {
    std::vector<std::complex<double>> v(n);
    to_c(data, v.data(), n);
    fft1(v.data(), n, sign);
    from_c(v.data(), data, n);
  }
}

void orc_fft2(int n0, int n1, int sign, int prec, double* data) {
  size_t n = (size_t)n0 * n1;
  if (prec == 32) {
    std::vector<std::complex<float>> v(n);
    to_c(data, v.data(), n);
    fft2(v.data(), n0, n1, sign);
    from_c(v.data(), data, n);
  } else {
    std::vector<std::complex<double>> v(n);
    to_c(data, v.data(), n);
    fft2(v.data(), n0, n1, sign);
    from_c(v.data(), data, n);
  }
}

// Exact NDFT: f_j = sum_{X,Y} img[X,Y] e^{-2 pi i (x_j (X - N0/2) + y_j (Y - N1/2))}
// (the operator the forward NUFFT approximates), all in double.
void orc_ndft_forward(int N0, int N1, int M, const float* nodes, const double* img, double* f) {
  for (int j = 0; j < M; j++) {
    double xj = nodes[2 * j], yj = nodes[2 * j + 1];
    double sr = 0, si = 0;
    for (int X = 0; X < N0; X++) {
      for (int Y = 0; Y < N1; Y++) {
        double ph = -2.0 * M_PI * (xj * (X - N0 / 2) + yj * (Y - N1 / 2));
        double c = std::cos(ph), s = std::sin(ph);
        double ar = img[2 * ((size_t)X * N1 + Y)], ai = img[2 * ((size_t)X * N1 + Y) + 1];
        sr += ar * c - ai * s;
        si += ar * s + ai * c;
      }
    }
    f[2 * j] = sr; f[2 * j + 1] = si;
  }
}

// Exact adjoint NDFT: img[X,Y] = sum_j f_j e^{+2 pi i (x_j (X - N0/2) + y_j (Y - N1/2))}.
void orc_ndft_adjoint(int N0, int N1, int M, const float* nodes, const double* f, double* img) {
  for (int X = 0; X < N0; X++) {
    for (int Y = 0; Y < N1; Y++) {
      double sr = 0, si = 0;
      for (int j = 0; j < M; j++) {
        double ph = 2.0 * M_PI * ((double)nodes[2 * j] * (X - N0 / 2) + (double)nodes[2 * j + 1] * (Y - N1 / 2));
        double c = std::cos(ph), s = std::sin(ph);
        sr += f[2 * j] * c - f[2 * j + 1] * s;
        si += f[2 * j] * s + f[2 * j + 1] * c;
      }
      img[2 * ((size_t)X * N1 + Y)] = sr;
      img[2 * ((size_t)X * N1 + Y) + 1] = si;
    }
  }
}

// One slice solve.  resid: iters+1 doubles or NULL; alpha: iters or NULL;
// m_init: N0*N1 warm start or NULL (zeros, as the reference).
int orc_icon_solve(void* h, const float* sino, int pre, int icon, int post, int dataType,
                   float thr, float* m_out, double* resid, double* alpha, const float* m_init) {
  auto* a = (AnyPlan*)h;
  SolveOpts o{pre, icon, post, dataType, thr};
  if (a->p32) icon_solve(*a->p32, sino, o, m_out, resid, alpha, m_init);
  else icon_solve(*a->p64, sino, o, m_out, resid, alpha, m_init);
  return 0;
}

// S independent slices (sinos: S x nz x N0, m_out: S x N0 x N1), OpenMP over
// slices with `threads` threads (<= 0: library default).  Returns the thread
// count used.
int orc_icon_solve_batch(void* h, int S, const float* sinos, int pre, int icon, int post,
                         int dataType, float thr, float* m_out, double* resid, int threads) {
  auto* a = (AnyPlan*)h;
  SolveOpts o{pre, icon, post, dataType, thr};
  int M = a->p32 ? a->p32->M : a->p64->M;
  int NN = a->p32 ? a->p32->N0 * a->p32->N1 : a->p64->N0 * a->p64->N1;
  int iters = pre + icon + post;
  int used = 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
  }
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int s = 0; s < S; s++) {
    const float* sino = sinos + (size_t)s * M;
    float* mo = m_out + (size_t)s * NN;
    double* rs = resid ? resid + (size_t)s * (iters + 1) : nullptr;
    if (a->p32) icon_solve(*a->p32, sino, o, mo, rs, nullptr, nullptr);
    else icon_solve(*a->p64, sino, o, mo, rs, nullptr, nullptr);
  }
  (void)threads;
  return used;
}

// ICONPreProcess (sourceCode/ICONPreProcess/ICONPreProcess.c:31-86, 256-302),
// per section k of a stack [nz][ny][nx]:
//   minusmode: min/max over the central region (10 % margins excluded, in
//   double), a 256-bin histogram of that region with bin floor((v-min)/step),
//   step = (max-min)/256; the reference's index 256 (v == max) lands past its
//   array and is never read, so it is not counted here; mode = first bin of
//   the largest count * step + min; v -= mode over the whole section.
//   normalize (tilts != NULL and thickness > 0): mean = sum (float)(v/size),
//   var = sum (v-mean)^2/size (double, in index order), v = v/sqrt(var) *
//   sqrt(0.33 thickness / cos(theta/180*3.1415926)), theta the negated angle.
// Parity: pinned by closed forms (tests/test_oracle.py); the reference
// program needs generated code to build.
void orc_preprocess(float* stack, int nx, int ny, int nz, const float* tilt_deg, int thickness, double* modes) {
  const int xe = (int)(nx * 0.1), ye = (int)(ny * 0.1);
  const size_t size = (size_t)nx * ny;
  for (int k = 0; k < nz; k++) {
    float* in = stack + (size_t)k * size;
    double mn = in[(size_t)ye * nx + xe], mx = mn;
    for (int j = ye; j < ny - ye; j++)
      for (int i = xe; i < nx - xe; i++) {
        const double v = in[(size_t)j * nx + i];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
      }
    const double step = (mx - mn) / 256.0;
    double mode = mn;
    if (step > 0) {
      std::vector<int> hist(256, 0);
      for (int j = ye; j < ny - ye; j++)
        for (int i = xe; i < nx - xe; i++) {
          const int b = (int)std::floor(((double)in[(size_t)j * nx + i] - mn) / step);
          if (b >= 0 && b < 256) hist[b]++;
        }
      int best = -1000, bi = 0;
      for (int b = 0; b < 256; b++)
        if (best < hist[b]) { best = hist[b]; bi = b; }
      mode = bi * step + mn;
    }
    if (modes) modes[k] = mode;
    for (size_t i = 0; i < size; i++) in[i] = (float)((double)in[i] - mode);
    if (tilt_deg && thickness > 0) {
      const double angle = -(double)tilt_deg[k];
      const double factor = 0.33 * thickness / std::cos(angle / 180.0 * 3.1415926);
      double sum = 0;
      for (size_t i = 0; i < size; i++) sum += (double)(in[i] / (float)size);   // float division, as written
      double var = 0;
      for (size_t i = 0; i < size; i++) var += ((double)in[i] - sum) * ((double)in[i] - sum) / (double)size;
      for (size_t i = 0; i < size; i++) in[i] = (float)((double)in[i] / std::sqrt(var) * std::sqrt(factor));
    }
  }
}

void orc_reproject(void* h, const float* m, float* out) {
  auto* a = (AnyPlan*)h;
  if (a->p32) reproject(*a->p32, m, out); else reproject(*a->p64, m, out);
}

}  // extern "C"
