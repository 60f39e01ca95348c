// All-FP64 ICON solve on the device: the reference solve in double precision
// (the reference's double build, DATATYPE = double, nufft_gpu_v8.cuh:25-33),
// every table, transform, spread, interpolation, reduction and update in
// double.  It is the device-side precision reference of the FP32 solve, not a
// fast path: plain kernels, one block per FFT line in LDS, the spread as FP64
// atomics, launches issued eagerly.
//
// Algorithm (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
//   Fourier-slice data   readSlicePart.cu:10-22, nufft_gpu_v8.cu:1432-1516
//   forward NUFFT        nufft_gpu_v8.cu:754-809 (deapodise + pad), 2D FFT,
//                        :669-727 (14 x 14 window taps, two-phase sum)
//   adjoint NUFFT        :1078-1102 (spread), inverse 2D FFT, :1104-1121 (crop)
//   iteration            Iteration2.cu:33-146, alpha_new2.cu:28-114,
//                        ICON/INFR update :1597-1639
//   reprojection         reProject_NFFT.cu:81-98 (1D adjoint at the omit tilt)
// Window: the reference's Gaussian, m = 6, with sigma = n/N, b and the taps
// computed in double from the float node coordinates of px2D.cu:8-47.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "icon_plan.h"

using namespace icon;

namespace {

constexpr int kM = 6;
constexpr int kT = 2 * kM + 2;   // 14 taps

__device__ __forceinline__ double2 d2(double x, double y) { return make_double2(x, y); }

// ---------------------------------------------------------------- kernels

// grid[s][x][y] = img[s][X][Y] / ck[X][Y] for X = (x + N0/2) mod n0 < N0 (and
// likewise Y), zero elsewhere; img real (m) or complex (rhat).
__global__ void k64_pad(const double* m, const double2* z, const double* ck, double2* grid, int S, int N0, int N1,
                        int n0, int n1) {
  const long total = (long)S * n0 * n1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int y = (int)(i % n1);
    const int x = (int)((i / n1) % n0);
    const int s = (int)(i / ((long)n0 * n1));
    const int X = (x + N0 / 2) % n0, Y = (y + N1 / 2) % n1;
    double2 v = d2(0.0, 0.0);
    if (X < N0 && Y < N1) {
      const long q = ((long)s * N0 + X) * N1 + Y;
      const double c = ck[(long)X * N1 + Y];
      v = z ? d2(z[q].x / c, z[q].y / c) : d2(m[q] / c, 0.0);
    }
    grid[i] = v;
  }
}

// img[s][X][Y] = grid[s][(X + n0 - N0/2) mod n0][(Y + n1 - N1/2) mod n1] / ck
__global__ void k64_crop(const double2* grid, const double* ck, double2* img, int S, int N0, int N1, int n0, int n1) {
  const long total = (long)S * N0 * N1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int Y = (int)(i % N1);
    const int X = (int)((i / N1) % N0);
    const int s = (int)(i / ((long)N0 * N1));
    const int x = (X + n0 - N0 / 2) % n0, y = (Y + n1 - N1 / 2) % n1;
    const double c = ck[(long)X * N1 + Y];
    const double2 v = grid[((long)s * n0 + x) * n1 + y];
    img[i] = d2(v.x / c, v.y / c);
  }
}

// Unnormalised radix-2 transform of `lines` lines of length n (power of two,
// <= 2048), one block per line in LDS: element e of line L sits at
// (L / lps) * sstride + (L % lps) * ls + e * es.  tw[q] = exp(2 pi i q / n),
// q < n/2; sign -1 uses its conjugate.
__global__ __launch_bounds__(1024) void k64_fft(double2* data, const double2* tw, int n, int lines, int lps,
                                                long sstride, long ls, long es, int sign) {
  extern __shared__ double2 a[];
  const int L = blockIdx.x;
  if (L >= lines) return;
  double2* base = data + (long)(L / lps) * sstride + (long)(L % lps) * ls;
  int logn = 0;
  while ((1 << logn) < n) logn++;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int r = (int)(__brev((unsigned)e) >> (32 - logn));
    a[r] = base[(long)e * es];
  }
  __syncthreads();
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const int k = b % half;
      const int i = (b / half) * len + k, j = i + half;
      const double2 w = tw[k * step];
      const double wy = sign < 0 ? -w.y : w.y;
      const double2 u = a[i], v = a[j];
      const double vr = v.x * w.x - v.y * wy, vi = v.x * wy + v.y * w.x;
      a[i] = d2(u.x + vr, u.y + vi);
      a[j] = d2(u.x - vr, u.y - vi);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < n; e += blockDim.x) base[(long)e * es] = a[e];
}

// f[s][j] = sum over the 14 x 14 taps of grid * phix * phiy: for every y-tap
// the 14 x-taps are summed with fused multiply-adds, then the 14 row sums in
// order (nufft_gpu_v8.cu:669-727).
__global__ void k64_interp(const double2* grid, const double* phix, const double* phiy, const int* ix, const int* iy,
                           double2* f, int S, int M, int n0, int n1) {
  const long total = (long)S * M;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int j = (int)(i % M), s = (int)(i / M);
    const double2* g = grid + (long)s * n0 * n1;
    double fx = 0.0, fy = 0.0;
    for (int ty = 0; ty < kT; ty++) {
      const int y = iy[(long)j * kT + ty];
      const double py = phiy[(long)j * kT + ty];
      double ax = 0.0, ay = 0.0;
      for (int tx = 0; tx < kT; tx++) {
        const double phi = phix[(long)j * kT + tx] * py;
        const double2 gv = g[(long)ix[(long)j * kT + tx] * n1 + y];
        ax = fma(gv.x, phi, ax);
        ay = fma(gv.y, phi, ay);
      }
      fx += ax;
      fy += ay;
    }
    f[i] = d2(fx, fy);
  }
}

// grid += f[s][j] * phix * phiy over every node's taps (FP64 atomics: the sum
// order per cell is not fixed, the values agree to rounding).
__global__ void k64_spread(const double2* f, const double* phix, const double* phiy, const int* ix, const int* iy,
                           double2* grid, int S, int M, int n0, int n1) {
  const long total = (long)S * M;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int j = (int)(i % M), s = (int)(i / M);
    double2* g = grid + (long)s * n0 * n1;
    const double2 fj = f[i];
    for (int ty = 0; ty < kT; ty++) {
      const int y = iy[(long)j * kT + ty];
      const double py = phiy[(long)j * kT + ty];
      for (int tx = 0; tx < kT; tx++) {
        const double phi = phix[(long)j * kT + tx] * py;
        double* c = reinterpret_cast<double*>(g + (long)ix[(long)j * kT + tx] * n1 + y);
        atomicAdd(c, fj.x * phi);
        atomicAdd(c + 1, fj.y * phi);
      }
    }
  }
}

__global__ void k64_zero(double* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

__device__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// One block (1024 threads) per slice.
// mode 0: out[s] = sum_j w_j |f_j - b_j|^2 (b given) or sum_j w_j |f_j|^2 (b null);
//         fw = w f when fw is given
// mode 1: rhat = img - ahwb, out[s] = sum |rhat|^2
__global__ __launch_bounds__(1024) void k64_reduce(int mode, const double2* f, const double2* b, const float* w,
                                                   double2* fw, int M, const double2* img, const double2* ahwb,
                                                   double2* rhat, long NN, double* out) {
  __shared__ double red[1024];
  const int s = blockIdx.x;
  double acc = 0.0;
  if (mode == 0) {
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
      const double2 v = f[(long)s * M + j];
      const double wj = (double)w[j];
      double dx = v.x, dy = v.y;
      if (b) {
        dx -= b[(long)s * M + j].x;
        dy -= b[(long)s * M + j].y;
      }
      acc += wj * (dx * dx + dy * dy);
      if (fw) fw[(long)s * M + j] = d2(v.x * wj, v.y * wj);
    }
  } else {
    for (long q = threadIdx.x; q < NN; q += blockDim.x) {
      const double2 a = img[(long)s * NN + q], c = ahwb[(long)s * NN + q];
      const double2 r = d2(a.x - c.x, a.y - c.y);
      rhat[(long)s * NN + q] = r;
      acc += r.x * r.x + r.y * r.y;
    }
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[s] = t;
}

// m <- m - alpha Re(rhat), then the ICON clamp (phase 2): dataType 1 m <= thr,
// dataType 2 m >= thr.
__global__ void k64_update(double* m, const double2* rhat, const double* alpha, long NN, int S, int phase,
                           int data_type, double thr) {
  const long total = (long)S * NN;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    double v = fma(-alpha[i / NN], rhat[i].x, m[i]);
    if (phase == 2) {
      if (data_type == 1) { if (v > thr) v = thr; }
      else { if (v < thr) v = thr; }
    }
    m[i] = v;
  }
}

__global__ void k64_alpha(const double* rtr, const double* rtar, double* alpha, double* trace, int stride, int idx,
                          int S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const double al = rtar[s] != 0.0 ? rtr[s] / rtar[s] : 0.0;
  alpha[s] = al;
  if (trace) trace[(long)s * stride + idx] = al;
}

__global__ void k64_resid(const double* num, const double* bnorm, double* resid, int stride, int idx, int S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  resid[(long)s * stride + idx] = bnorm[s] > 0.0 ? sqrt(num[s] / bnorm[s]) : 0.0;
}

// 1D rows (projection rows or the omit-tilt line): row r of length N0, to and
// from lines of length n0.
// pad: g[r][x] = in[r][X] / ck1[X], X = (x + N0/2) mod n0 < N0 (real input from
// the float sinogram, or the cross-validation row map: input row of node row
// r = s nz + i is s nz_in + i)
__global__ void k64_pad1(const float* sino, const double* ck1, double2* g, int rows, int N0, int n0) {
  const long total = (long)rows * n0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % n0), r = (int)(i / n0);
    const int X = (x + N0 / 2) % n0;
    g[i] = X < N0 ? d2((double)sino[(long)r * N0 + X] / ck1[X], 0.0) : d2(0.0, 0.0);
  }
}

// b[r][k] = sum_t g[r][i1[k][t]] phi1[k][t] (fused multiply-adds in tap order)
__global__ void k64_gather1(const double2* g, const double* phi1, const int* i1, double2* b, int rows, int N0,
                            int n0) {
  const long total = (long)rows * N0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % N0), r = (int)(i / N0);
    double fx = 0.0, fy = 0.0;
    for (int t = 0; t < kT; t++) {
      const double ph = phi1[(long)k * kT + t];
      const double2 gv = g[(long)r * n0 + i1[(long)k * kT + t]];
      fx = fma(gv.x, ph, fx);
      fy = fma(gv.y, ph, fy);
    }
    b[i] = d2(fx, fy);
  }
}

// the 1D adjoint of the omit-tilt values of every slice, one thread per
// (slice, cell): the nodes touching a cell summed in ascending node order
__global__ void k64_spread1(const double2* f, int M, int iomit, const double* phi1, const int* i1, double2* g, int S,
                            int N0, int n0) {
  const long total = (long)S * n0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % n0), s = (int)(i / n0);
    const double2* fs = f + (long)s * M + (long)iomit * N0;
    double cx = 0.0, cy = 0.0;
    for (int k = 0; k < N0; k++)
      for (int t = 0; t < kT; t++)
        if (i1[(long)k * kT + t] == x) {
          const double ph = phi1[(long)k * kT + t];
          cx = fma(fs[k].x, ph, cx);
          cy = fma(fs[k].y, ph, cy);
        }
    g[i] = d2(cx, cy);
  }
}

__global__ void k64_crop1(const double2* g, const double* ck1, double* out, int S, int N0, int n0) {
  const long total = (long)S * N0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int X = (int)(i % N0), s = (int)(i / N0);
    const int x = (X + n0 - N0 / 2) % n0;
    out[i] = g[(long)s * n0 + x].x / ck1[X];
  }
}

// ---------------------------------------------------------------- host

// Window taps of node coordinates c (float) on a grid of n: l = floor(n c) - m + t,
// phi = bf exp(-(n (c - l/n))^2 / b), index l mod n (nufft_gpu_v8.cu:221-235 in double).
void taps64(const float* c, int stride, int M, int n, double b, double bf, std::vector<double>& phi,
            std::vector<int>& idx) {
  phi.resize((size_t)M * kT);
  idx.resize((size_t)M * kT);
  for (int j = 0; j < M; j++) {
    const double x = c[(size_t)j * stride];
    for (int t = 0; t < kT; t++) {
      const int l = (int)std::floor(n * x) - kM + t;
      const double nx1 = n * (x - (double)l / n);
      phi[(size_t)j * kT + t] = bf * std::exp(-(nx1 * nx1 / b));
      idx[(size_t)j * kT + t] = ((l % n) + n) % n;
    }
  }
}

// exp(-b (pi k / n)^2), the deapodisation factor of one dimension
double ck64(int k, int n, double b) {
  const double a = M_PI * (double)k / (double)n;
  return std::exp(-b * a * a);
}

struct Dev {
  std::vector<void*> ptrs;
  ~Dev() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  T* alloc(size_t n, hipError_t* e) {
    void* p = nullptr;
    if (*e == hipSuccess) *e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
    if (*e == hipSuccess) ptrs.push_back(p);
    return (T*)p;
  }
  template <class T>
  T* upload(const std::vector<T>& h, hipError_t* e) {
    T* p = alloc<T>(h.size(), e);
    if (*e == hipSuccess && !h.empty()) *e = hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    return p;
  }
};

int blocks_for(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

int solve_fp64(int nx, int thickness, int nz, const float* tilt_deg, int S, const float* sino, double* m_out,
               double* reproj, double* resid, double* alpha, const icon_solve_opts* opts);

}  // namespace

extern "C" int icon_reconstruct_fp64(int device, int nx, int thickness, int nz, const float* tilt_deg, int S,
                                     const float* sino, double* m_out, double* reproj, double* resid, double* alpha,
                                     const icon_solve_opts* opts) {
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (hipSetDevice(device) != hipSuccess) {
    set_last_error("icon_reconstruct_fp64: no such device");
    return ICON_ERR_ARG;
  }
  const int rc = solve_fp64(nx, thickness, nz, tilt_deg, S, sino, m_out, reproj, resid, alpha, opts);
  if (prev >= 0) (void)hipSetDevice(prev);
  return rc;
}

namespace {

// The solve on the current device; its buffers are freed on return.
int solve_fp64(int nx, int thickness, int nz, const float* tilt_deg, int S, const float* sino, double* m_out,
               double* reproj, double* resid, double* alpha, const icon_solve_opts* opts) {
  if (!tilt_deg || !sino || !m_out || !opts || S < 1) {
    set_last_error("icon_reconstruct_fp64: bad argument");
    return ICON_ERR_ARG;
  }
  const icon_solve_opts& o = *opts;
  if (o.pre < 0 || o.icon < 0 || o.post < 0 || (o.data_type != 1 && o.data_type != 2)) {
    set_last_error("icon_reconstruct_fp64: bad iteration split or data type");
    return ICON_ERR_ARG;
  }
  Geometry g;
  std::string err;
  if (!build_geometry(g, nx, thickness > 0 ? thickness : nx, nz, tilt_deg, &err)) {
    set_last_error(err.c_str());
    return ICON_ERR_ARG;
  }
  const int N0 = g.N0, N1 = g.N1, M = g.M, n0 = g.n0, n1 = g.n1;
  // FP64 tables: sigma = n/N, b = 2 sigma m / ((2 sigma - 1) pi), bf = 1/sqrt(pi b)
  const double s0 = (double)n0 / N0, s1 = (double)n1 / N1;
  const double b0 = 2 * s0 * kM / ((2 * s0 - 1) * M_PI), b1 = 2 * s1 * kM / ((2 * s1 - 1) * M_PI);
  const double bf0 = 1.0 / std::sqrt(M_PI * b0), bf1 = 1.0 / std::sqrt(M_PI * b1);
  std::vector<double> phix, phiy, phi1, ck((size_t)N0 * N1), ck1(N0);
  std::vector<int> ix, iy, i1;
  taps64(g.nodes.data(), 2, M, n0, b0, bf0, phix, ix);
  taps64(g.nodes.data() + 1, 2, M, n1, b1, bf1, phiy, iy);
  std::vector<float> px1(N0);
  for (int k = 0; k < N0; k++) px1[k] = (float)(((double)(float)k - (double)(float)N0 / 2.0) / (double)(float)N0);
  taps64(px1.data(), 1, N0, n0, b0, bf0, phi1, i1);
  for (int X = 0; X < N0; X++) {
    ck1[X] = ck64(X - N0 / 2, n0, b0);
    for (int Y = 0; Y < N1; Y++) ck[(size_t)X * N1 + Y] = ck64(X - N0 / 2, n0, b0) * ck64(Y - N1 / 2, n1, b1);
  }
  auto twid = [](int n) {
    std::vector<double2> t(n / 2);
    for (int q = 0; q < n / 2; q++) t[q] = make_double2(std::cos(2 * M_PI * q / n), std::sin(2 * M_PI * q / n));
    return t;
  };
  hipError_t e = hipSuccess;
  Dev d;
  const size_t NN = (size_t)N0 * N1, G = (size_t)n0 * n1;
  double* d_phix = d.upload(phix, &e);
  double* d_phiy = d.upload(phiy, &e);
  int* d_ix = d.upload(ix, &e);
  int* d_iy = d.upload(iy, &e);
  double* d_phi1 = d.upload(phi1, &e);
  int* d_i1 = d.upload(i1, &e);
  double* d_ck = d.upload(ck, &e);
  double* d_ck1 = d.upload(ck1, &e);
  float* d_w = d.upload(g.w, &e);
  double2* d_tw0 = d.upload(twid(n0), &e);
  double2* d_tw1 = d.upload(twid(n1), &e);
  float* d_sino = d.alloc<float>((size_t)S * nz * N0, &e);
  double2* d_grid = d.alloc<double2>((size_t)S * G, &e);
  double2* d_b = d.alloc<double2>((size_t)S * M, &e);
  double2* d_f = d.alloc<double2>((size_t)S * M, &e);
  double2* d_fw = d.alloc<double2>((size_t)S * M, &e);
  double2* d_img = d.alloc<double2>((size_t)S * NN, &e);
  double2* d_ahwb = d.alloc<double2>((size_t)S * NN, &e);
  double2* d_rhat = d.alloc<double2>((size_t)S * NN, &e);
  double* d_m = d.alloc<double>((size_t)S * NN, &e);
  double* d_bnorm = d.alloc<double>(S, &e);
  double* d_num = d.alloc<double>(S, &e);
  double* d_rtr = d.alloc<double>(S, &e);
  double* d_al = d.alloc<double>(S, &e);
  const int iters = o.pre + o.icon + o.post;
  double* d_res = d.alloc<double>((size_t)S * (iters + 1), &e);
  double* d_altr = d.alloc<double>((size_t)S * (iters > 0 ? iters : 1), &e);
  double2* d_g1 = d.alloc<double2>((size_t)S * nz * n0, &e);
  double* d_rp = d.alloc<double>((size_t)S * N0, &e);
  if (e == hipSuccess) e = hipMemcpy(d_sino, sino, sizeof(float) * S * nz * N0, hipMemcpyHostToDevice);
  auto fft2 = [&](int sign) {
    // y (rows of n1), then x (columns of n0)
    hipLaunchKernelGGL(k64_fft, dim3(S * n0), dim3(std::min(n1 / 2, 1024)), sizeof(double2) * n1, 0, d_grid, d_tw1,
                       n1, S * n0, n0, (long)G, (long)n1, 1L, sign);
    hipLaunchKernelGGL(k64_fft, dim3(S * n1), dim3(std::min(n0 / 2, 1024)), sizeof(double2) * n0, 0, d_grid, d_tw0,
                       n0, S * n1, n1, (long)G, 1L, (long)n1, sign);
  };
  auto forward = [&](const double* m, const double2* z) {
    hipLaunchKernelGGL(k64_pad, dim3(blocks_for((long)S * G)), dim3(256), 0, 0, m, z, d_ck, d_grid, S, N0, N1, n0, n1);
    fft2(-1);
    hipLaunchKernelGGL(k64_interp, dim3(blocks_for((long)S * M)), dim3(256), 0, 0, d_grid, d_phix, d_phiy, d_ix, d_iy,
                       d_f, S, M, n0, n1);
  };
  auto adjoint = [&](const double2* f, double2* img) {
    hipLaunchKernelGGL(k64_zero, dim3(blocks_for((long)S * G * 2)), dim3(256), 0, 0, (double*)d_grid, (long)S * G * 2);
    hipLaunchKernelGGL(k64_spread, dim3(blocks_for((long)S * M)), dim3(256), 0, 0, f, d_phix, d_phiy, d_ix, d_iy,
                       d_grid, S, M, n0, n1);
    fft2(1);
    hipLaunchKernelGGL(k64_crop, dim3(blocks_for((long)S * NN)), dim3(256), 0, 0, d_grid, d_ck, img, S, N0, N1, n0, n1);
  };
  if (e == hipSuccess) {
    // b: the 1D transform of every projection row (cross-validation plans are not offered here)
    const int rows = S * nz;
    hipLaunchKernelGGL(k64_pad1, dim3(blocks_for((long)rows * n0)), dim3(256), 0, 0, d_sino, d_ck1, d_g1, rows, N0, n0);
    hipLaunchKernelGGL(k64_fft, dim3(rows), dim3(std::min(n0 / 2, 1024)), sizeof(double2) * n0, 0, d_g1, d_tw0, n0,
                       rows, 1, (long)n0, 0L, 1L, -1);
    hipLaunchKernelGGL(k64_gather1, dim3(blocks_for((long)rows * N0)), dim3(256), 0, 0, d_g1, d_phi1, d_i1, d_b, rows,
                       N0, n0);
    // bnorm = sum w |b|^2, fw = w b, A^H W b
    hipLaunchKernelGGL(k64_reduce, dim3(S), dim3(1024), 0, 0, 0, d_b, (const double2*)nullptr, d_w, d_fw, M,
                       (const double2*)nullptr, (const double2*)nullptr, (double2*)nullptr, 0L, d_bnorm);
    adjoint(d_fw, d_ahwb);
    hipLaunchKernelGGL(k64_zero, dim3(blocks_for((long)S * NN)), dim3(256), 0, 0, d_m, (long)(S * NN));
    for (int it = 0; it < iters; it++) {
      forward(d_m, nullptr);
      hipLaunchKernelGGL(k64_reduce, dim3(S), dim3(1024), 0, 0, 0, d_f, d_b, d_w, d_fw, M, (const double2*)nullptr,
                         (const double2*)nullptr, (double2*)nullptr, 0L, d_num);
      hipLaunchKernelGGL(k64_resid, dim3((S + 63) / 64), dim3(64), 0, 0, d_num, d_bnorm, d_res, iters + 1, it, S);
      adjoint(d_fw, d_img);
      hipLaunchKernelGGL(k64_reduce, dim3(S), dim3(1024), 0, 0, 1, (const double2*)nullptr, (const double2*)nullptr,
                         d_w, (double2*)nullptr, 0, d_img, d_ahwb, d_rhat, (long)NN, d_rtr);
      forward(nullptr, d_rhat);
      hipLaunchKernelGGL(k64_reduce, dim3(S), dim3(1024), 0, 0, 0, d_f, (const double2*)nullptr, d_w,
                         (double2*)nullptr, M, (const double2*)nullptr, (const double2*)nullptr, (double2*)nullptr,
                         0L, d_num);
      hipLaunchKernelGGL(k64_alpha, dim3((S + 63) / 64), dim3(64), 0, 0, d_rtr, d_num, d_al, d_altr, iters, it, S);
      const int phase = (o.pre <= it && it < iters - o.post) ? 2 : 1;
      hipLaunchKernelGGL(k64_update, dim3(blocks_for((long)S * NN)), dim3(256), 0, 0, d_m, d_rhat, d_al, (long)NN, S,
                         phase, o.data_type, (double)o.threshold);
    }
    // R of the result and the reprojection at the omit tilt
    forward(d_m, nullptr);
    hipLaunchKernelGGL(k64_reduce, dim3(S), dim3(1024), 0, 0, 0, d_f, d_b, d_w, (double2*)nullptr, M,
                       (const double2*)nullptr, (const double2*)nullptr, (double2*)nullptr, 0L, d_num);
    hipLaunchKernelGGL(k64_resid, dim3((S + 63) / 64), dim3(64), 0, 0, d_num, d_bnorm, d_res, iters + 1, iters, S);
    hipLaunchKernelGGL(k64_spread1, dim3(blocks_for((long)S * n0)), dim3(256), 0, 0, d_f, M, g.iomit, d_phi1, d_i1,
                       d_g1, S, N0, n0);
    hipLaunchKernelGGL(k64_fft, dim3(S), dim3(std::min(n0 / 2, 1024)), sizeof(double2) * n0, 0, d_g1, d_tw0, n0, S, 1,
                       (long)n0, 0L, 1L, 1);
    hipLaunchKernelGGL(k64_crop1, dim3(blocks_for((long)S * N0)), dim3(256), 0, 0, d_g1, d_ck1, d_rp, S, N0, n0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(m_out, d_m, sizeof(double) * S * NN, hipMemcpyDeviceToHost);
  if (e == hipSuccess && reproj) e = hipMemcpy(reproj, d_rp, sizeof(double) * S * N0, hipMemcpyDeviceToHost);
  if (e == hipSuccess && resid) e = hipMemcpy(resid, d_res, sizeof(double) * S * (iters + 1), hipMemcpyDeviceToHost);
  if (e == hipSuccess && alpha && iters > 0)
    e = hipMemcpy(alpha, d_altr, sizeof(double) * S * iters, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_last_error((std::string("icon_reconstruct_fp64: ") + hipGetErrorString(e)).c_str());
    return ICON_ERR_DEVICE;
  }
  return ICON_OK;
}

}  // namespace
