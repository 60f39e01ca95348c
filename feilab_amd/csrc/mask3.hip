// ICONMask3 (sourceCode/ICONMask3/ICONCrossVMaskIn2D_weightOnlyOnSampled.cu):
// combine the reconstructed slices midYYYYY.mrc into a volume, each slice
// filtered in 2D Fourier space by
//   realMask  1 inside the FRC-0.5 radius; outside, the angular sampling
//             mask (the union of the tilt lines), raised to a Gaussian edge
//             between the FRC-0.5 and FRC-0.3 radii;
//   weighting |k| / r_full on sampled frequencies beyond r_full (the radius
//             where neighbouring tilt lines stop overlapping), else 1;
// then cropped to the specimen thickness.
//
// The masks are built on the host exactly as the reference builds them
// (float/double placement included).  The filtering runs batched on the
// device with this library's own 2D FFT: for a real slice f with spectrum F,
// the reference's R2C -> (half-spectrum mask M) -> C2R / N^2 equals
// Re(IFFT2(F * Mfull)) / N^2 with Mfull(ky, kx) = M(ky, kx) for kx <= N/2
// and M(-ky, N-kx) otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>
#include <vector>

#include "../../include/icon_mi355x.h"
#include "mrc_io.h"

namespace {

constexpr double kPi180 = 0.01745329252;   // PI_180 of the reference

thread_local std::string g_merr;

int merr(int code, const std::string& m) {
  g_merr = m;
  return code;
}

// The angular sampling mask (main, lines 958-1000): the central column of an
// N x N image rotated by -theta-90 degrees for every tilt with bilinear
// splatting, summed and clipped to 1.  Index j1*N + i1, centred at N/2.
void sampling_mask(int N, const float* thita, int ang_num, std::vector<float>& origmask) {
  const long nx = N, ny = N;
  std::vector<float> line(nx * ny, 0.f), rot(nx * ny);
  origmask.assign(nx * ny, 0.f);
  for (long j = 0; j < nx; j++) line[j * nx + nx / 2] = 1.f;
  for (int k = 0; k < ang_num; k++) {
    const float c = (float)std::cos((-thita[k] - 90) * kPi180);
    const float s = (float)std::sin((-thita[k] - 90) * kPi180);
    std::fill(rot.begin(), rot.end(), 0.f);
    for (long j = 0; j < ny; j++)
      for (long i = 0; i < nx; i++) {
        const float lv = line[j * nx + i];
        if (lv == 0.f) continue;
        const float xf = (float)(i - nx / 2) * c - (float)(j - ny / 2) * s + (float)(nx / 2);
        const float yf = (float)(i - nx / 2) * s + (float)(j - ny / 2) * c + (float)(ny / 2);
        const int xi = (int)xf, yi = (int)yf;
        const float dx = xf - xi, dy = yf - yi;
        if (xi >= 0 && xi < nx && yi >= 0 && yi < ny) rot[yi * nx + xi] += lv * (1 - dx) * (1 - dy);
        if (xi >= 0 && xi < nx && yi + 1 >= 0 && yi + 1 < ny) rot[(yi + 1) * nx + xi] += lv * (1 - dx) * dy;
        if (xi + 1 >= 0 && xi + 1 < nx && yi >= 0 && yi < ny) rot[yi * nx + xi + 1] += lv * dx * (1 - dy);
        if (xi + 1 >= 0 && xi + 1 < nx && yi + 1 >= 0 && yi + 1 < ny) rot[(yi + 1) * nx + xi + 1] += lv * dx * dy;
      }
    for (long i = 0; i < nx * ny; i++) origmask[i] += rot[i];
  }
  for (auto& v : origmask)
    if (v > 1) v = 1;
}

__global__ void k_real_to_complex(const float* in, float2* out, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = make_float2(in[i], 0.f);
}

// F *= Mfull, the half-spectrum mask mirrored onto the other half
__global__ void k_apply_mask(float2* f, const float* half, int N, long count) {
  const int H = N / 2 + 1;
  const long n = count * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int kx = (int)(i % N);
    const int ky = (int)((i / N) % N);
    const float m = kx <= N / 2 ? half[(long)ky * H + kx] : half[(long)((N - ky) % N) * H + (N - kx)];
    const float2 v = f[i];
    f[i] = make_float2(v.x * m, v.y * m);
  }
}

// out = Re(f) / (N*N), as the reference divides after its C2R in double
__global__ void k_real_scaled(const float2* f, float* out, long n, double factor) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (float)((double)f[i].x / factor);
}

int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 65536); }

// Widths that are not powers of two: each 1D DFT of length N by Bluestein's
// chirp-z method on this library's power-of-two row FFT of length
// M >= 2N - 1.  With c_n = exp(sign i pi n^2 / N),
//   X_k = c_k sum_n (x_n c_n) conj(c_{k-n}),
// a circular convolution of length M.  A pass transforms the rows of
// [count][N][N] and stores the result transposed, so two passes give the 2D
// transform in the input's orientation.
struct Bluestein {
  int N = 0, M = 0;
  float2* chirp[2] = {nullptr, nullptr};   // c_n for sign -1, +1
  float2* bhat[2] = {nullptr, nullptr};    // FFT_M of the conj(c) kernel
};

// c_n, n^2 reduced mod 2N so the phase stays exact in double
std::vector<std::complex<double>> chirp_host(int N, int sign) {
  std::vector<std::complex<double>> c(N);
  for (long n = 0; n < N; n++) {
    const double ph = sign * M_PI * (double)((n * n) % (2L * N)) / (double)N;
    c[n] = std::complex<double>(std::cos(ph), std::sin(ph));
  }
  return c;
}

hipError_t bluestein_init(Bluestein& b, int N) {
  b.N = N;
  b.M = 128;
  while (b.M < 2 * N - 1) b.M *= 2;
  const int M = b.M;
  std::vector<std::complex<double>> tw(M);
  for (int k = 0; k < M; k++) tw[k] = std::polar(1.0, -2.0 * M_PI * k / M);
  for (int d = 0; d < 2; d++) {
    const auto c = chirp_host(N, d ? 1 : -1);
    std::vector<std::complex<double>> kern(M, 0.0), hat(M);
    kern[0] = std::conj(c[0]);
    for (int j = 1; j < N; j++) kern[j] = kern[M - j] = std::conj(c[j]);
    for (int k = 0; k < M; k++) {   // FP64 DFT, exponent sign -1 like the device forward FFT
      std::complex<double> acc = 0;
      for (int j = 0; j < M; j++)
        if (kern[j] != 0.0) acc += kern[j] * tw[(size_t)((long)j * k % M)];
      hat[k] = acc / (double)M;     // the 1/M of the inverse FFT folded in
    }
    std::vector<float2> cf(N), hf(M);
    for (int n = 0; n < N; n++) cf[n] = make_float2((float)c[n].real(), (float)c[n].imag());
    for (int k = 0; k < M; k++) hf[k] = make_float2((float)hat[k].real(), (float)hat[k].imag());
    hipError_t e = hipMalloc(&b.chirp[d], sizeof(float2) * N);
    if (e == hipSuccess) e = hipMalloc(&b.bhat[d], sizeof(float2) * M);
    if (e == hipSuccess) e = hipMemcpy(b.chirp[d], cf.data(), sizeof(float2) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.bhat[d], hf.data(), sizeof(float2) * M, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void bluestein_free(Bluestein& b) {
  for (int d = 0; d < 2; d++) {
    if (b.chirp[d]) (void)hipFree(b.chirp[d]);
    if (b.bhat[d]) (void)hipFree(b.bhat[d]);
  }
}

__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// work[r][m] = x[r][m] c_m for m < N, 0 up to M
__global__ void k_bs_load(const float2* x, const float2* chirp, float2* work, int N, int M, long rows) {
  const long n = rows * M;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int m = (int)(i % M);
    const long r = i / M;
    work[i] = m < N ? cmul(x[r * N + m], chirp[m]) : make_float2(0.f, 0.f);
  }
}

__global__ void k_bs_mul(float2* work, const float2* bhat, int M, long rows) {
  const long n = rows * M;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    work[i] = cmul(work[i], bhat[i % M]);
}

// y[s][k][r] = c_k conv[s][r][k]: the pass's result, transposed per slice
__global__ void k_bs_store_t(const float2* work, const float2* chirp, float2* y, int N, int M, long count) {
  const long n = count * N * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i % N);
    const int k = (int)((i / N) % N);
    const long s = i / ((long)N * N);
    y[i] = cmul(work[(s * N + r) * M + k], chirp[k]);
  }
}

int bluestein_2d(const Bluestein& b, float2* d_f, float2* work, int count, int sign, hipStream_t st) {
  const int d = sign > 0 ? 1 : 0;
  const long rows = (long)count * b.N;
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL(k_bs_load, dim3(grid_for(rows * b.M)), dim3(256), 0, st, d_f, b.chirp[d], work, b.N, b.M, rows);
    if (hipGetLastError() != hipSuccess) return ICON_ERR_DEVICE;
    int rc = icon_fft_rows(b.M, (int)rows, -1, work, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bs_mul, dim3(grid_for(rows * b.M)), dim3(256), 0, st, work, b.bhat[d], b.M, rows);
    if (hipGetLastError() != hipSuccess) return ICON_ERR_DEVICE;
    if ((rc = icon_fft_rows(b.M, (int)rows, 1, work, st))) return rc;
    hipLaunchKernelGGL(k_bs_store_t, dim3(grid_for(rows * b.N)), dim3(256), 0, st, work, b.chirp[d], d_f, b.N, b.M,
                       (long)count);
    if (hipGetLastError() != hipSuccess) return ICON_ERR_DEVICE;
  }
  return ICON_OK;
}

}  // namespace

extern "C" {

const char* icon_mask3_error(void) { return g_merr.c_str(); }

int icon_mask3_weights(int N, const float* tilt_deg, int ang_num, float radius05_rate, float radius03_rate,
                       int radius_given, int gauss_width, float mean_ang_step, float* half_mask) {
  if (N < 2 || !tilt_deg || ang_num < 1 || !half_mask) return merr(ICON_ERR_ARG, "icon_mask3_weights: bad argument");
  std::vector<float> origmask;
  sampling_mask(N, tilt_deg, ang_num, origmask);
  // genMask (lines 369-455)
  const int r_full = (int)std::ceil(2.0 / (mean_ang_step / 360.0 * 2 * 3.1415926));
  const int radius05 = (int)((float)N / 2.0 * radius05_rate);
  const int radius03 = (int)((float)N / 2.0 * radius03_rate);
  const int nx = N, ny = N;
  long gaussianlenMax;
  const int radius = radius05;
  if (!radius_given)
    gaussianlenMax = gauss_width >= 0 ? gauss_width : radius03 - radius05;
  else
    gaussianlenMax = gauss_width >= 0 ? gauss_width : (long)(10.0 * (N / 150.0));
  if (gaussianlenMax < 10.0 * (N / 150.0)) gaussianlenMax = (long)(10.0 * (N / 150.0));
  const long gaussianradius = radius + gaussianlenMax;
  const long gaussianlen = gaussianradius - radius;
  const double c = (double)gaussianlen / 3.0;
  const double e_end = std::exp(-(double)(gaussianlen * gaussianlen) / (2 * c * c));
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx / 2 + 1; i++) {
      const int i1 = (i + nx / 2) % nx, j1 = (j + ny / 2) % ny;
      const long rr = (long)std::sqrt((double)((i1 - nx / 2) * (i1 - nx / 2) + (j1 - ny / 2) * (j1 - ny / 2)));
      const float om = origmask[(size_t)j1 * nx + i1];
      float real;
      if (rr >= radius) {
        if (rr <= gaussianradius) {
          double w = om;
          const double xxx = (double)(rr - radius);
          const double w2 = (std::exp(-(xxx * xxx) / (2 * c * c)) - e_end) / (1 - e_end);
          w = w > w2 ? w : w2;
          real = (float)w;
        } else {
          real = om;
        }
      } else {
        real = 1.f;
      }
      const float weighting = (rr > r_full && std::fabs(om) > 0.00001) ? (float)rr / (float)r_full : 1.f;
      half_mask[(size_t)j * (nx / 2 + 1) + i] = real * weighting;
    }
  return ICON_OK;
}

int icon_mask3_filter(int device, float* d_slices, int N, int count, const float* half_mask, void* stream) {
  if (!d_slices || !half_mask || count < 1 || N < 8 || N > 2048)
    return merr(ICON_ERR_ARG, "icon_mask3_filter: slices must be N x N with N in 8..2048");
  const bool pow2 = N >= 128 && (N & (N - 1)) == 0;
  if (!pow2 && N > 1024)
    return merr(ICON_ERR_ARG, "icon_mask3_filter: widths above 1024 must be powers of two");
  if (hipSetDevice(device) != hipSuccess) return merr(ICON_ERR_ARG, "icon_mask3_filter: no such device");
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)count * N * N;
  const size_t hb = sizeof(float) * (size_t)N * (N / 2 + 1);
  float2 *d_f = nullptr, *d_work = nullptr;
  float* d_m = nullptr;
  Bluestein bs;
  hipError_t e = hipMalloc(&d_f, sizeof(float2) * n);
  if (e == hipSuccess) e = hipMalloc(&d_m, hb);
  if (e == hipSuccess) e = hipMemcpy(d_m, half_mask, hb, hipMemcpyHostToDevice);
  if (e == hipSuccess && !pow2) e = bluestein_init(bs, N);
  if (e == hipSuccess && !pow2) e = hipMalloc(&d_work, sizeof(float2) * (size_t)count * N * bs.M);
  int rc = ICON_OK;
  auto fft2 = [&](int sign) {
    return pow2 ? icon_fft_2d(N, N, count, sign, d_f, stream) : bluestein_2d(bs, d_f, d_work, count, sign, st);
  };
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_real_to_complex, dim3(grid_for(n)), dim3(256), 0, st, d_slices, d_f, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) rc = fft2(-1);
  if (e == hipSuccess && rc == ICON_OK) {
    hipLaunchKernelGGL(k_apply_mask, dim3(grid_for(n)), dim3(256), 0, st, d_f, d_m, N, (long)count);
    e = hipGetLastError();
  }
  if (e == hipSuccess && rc == ICON_OK) rc = fft2(1);
  if (e == hipSuccess && rc == ICON_OK) {
    hipLaunchKernelGGL(k_real_scaled, dim3(grid_for(n)), dim3(256), 0, st, d_f, d_slices, n, (double)N * (double)N);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d_f) (void)hipFree(d_f);
  if (d_m) (void)hipFree(d_m);
  if (d_work) (void)hipFree(d_work);
  bluestein_free(bs);
  if (e != hipSuccess) return merr(ICON_ERR_DEVICE, std::string("icon_mask3_filter: ") + hipGetErrorString(e));
  if (rc == ICON_ERR_DEVICE) return merr(rc, "icon_mask3_filter: kernel launch failed");
  if (rc != ICON_OK) return merr(rc, std::string("icon_mask3_filter: ") + icon_last_error());
  return ICON_OK;
}

int icon_mask3(int device, const char* input_dir, int slice_begin, int slice_end, int thickness,
               const char* tilt_path, const char* out_path, int radius, const char* crossv_frc,
               const char* fullrec_frc, int zshift, int gauss_width, int no_mask) {
  if (!input_dir || !tilt_path || !out_path || slice_end < slice_begin || thickness < 1)
    return merr(ICON_ERR_ARG, "parameter error! '-inputPath' or '-slice' or '-thickness' or '-tiltfile' or "
                              "'-output' are missed!");
  auto mid = [&](int z) {
    char b[4096];
    std::snprintf(b, sizeof b, "%s/mid%05d.mrc", input_dir, z);
    return std::string(b);
  };
  // geometry of the slices
  MrcHeader head;
  {
    FILE* fp = std::fopen(mid(slice_begin).c_str(), "rb");
    if (!fp || !mrc::read_header(fp, &head)) {
      if (fp) std::fclose(fp);
      return merr(ICON_ERR_IO, "Can not open infile " + mid(slice_begin) + "!");
    }
    std::fclose(fp);
  }
  const int Sx = head.nx, Sy = head.ny;
  const int ystart = Sy / 2 - thickness / 2 + zshift;
  if (ystart < 0 || ystart + thickness > Sy)
    return merr(ICON_ERR_ARG, "zshift & thickness make the size of volume out of limitation!");
  const int nslice = slice_end - slice_begin + 1;
  std::vector<float> out((size_t)nslice * thickness * Sx);
  std::vector<float> half;
  if (!no_mask) {
    if (Sx != Sy)
      return merr(ICON_ERR_ARG, "icon_mask3: the filtering needs square slices (thickness of the ICON run = width)");
    // tilt file (main, lines 803-822): numbers split by " ;,\n\r\t", up to "END"
    std::vector<float> thita;
    {
      FILE* fa = std::fopen(tilt_path, "r");
      if (!fa) return merr(ICON_ERR_IO, std::string("Cannot open file ") + tilt_path);
      char line[800];
      bool stop = false;
      while (!stop && std::fgets(line, sizeof line, fa)) {
        char* p = std::strtok(line, " ;,\n\r\t");
        if (!p || strncasecmp(p, "END", 3) == 0) break;
        while (p) {
          thita.push_back((float)std::atof(p));
          p = std::strtok(nullptr, " ;,\n\r\t");
        }
      }
      std::fclose(fa);
    }
    if (thita.size() < 2) return merr(ICON_ERR_ARG, "icon_mask3: the tilt file needs at least two angles");
    float biggest = -1, smallest = 1;
    for (float t : thita) {
      if (biggest < t) biggest = t;
      if (smallest > t) smallest = t;
    }
    const float big_step = biggest - smallest;
    const int big_step_radius = (int)std::ceil(2 * 180.0 / 3.1415926 / big_step);
    float mean_step = 0.f;
    for (size_t i = 0; i + 1 < thita.size(); i++) mean_step += std::fabs(thita[i + 1] - thita[i]);
    mean_step /= (float)(thita.size() - 1);
    // radii from the FRC files (lines 866-940)
    bool radius_given = radius > 0;
    float r05 = 0.f, r03 = 0.f;
    if (!radius_given) {
      auto read_frc = [](const char* path, std::vector<float>& v) {
        FILE* f = path ? std::fopen(path, "r") : nullptr;
        if (!f) return false;
        float a, b;
        while (std::fscanf(f, "%f %f", &a, &b) == 2) v.push_back(b);
        std::fclose(f);
        return true;
      };
      std::vector<float> cv, fr;
      if (!read_frc(crossv_frc, cv)) return merr(ICON_ERR_IO, std::string("Cannot open file '") + (crossv_frc ? crossv_frc : "") + "'");
      bool valid = (int)cv.size() >= Sx / 2 - 1;
      if (valid) {
        if (!read_frc(fullrec_frc, fr)) return merr(ICON_ERR_IO, std::string("Cannot open file '") + (fullrec_frc ? fullrec_frc : "") + "'");
        valid = (int)fr.size() >= Sx / 2 - 1;
      }
      if (valid) {
        const int sz = (int)fr.size();
        for (int i = 0; i < sz && i < (int)cv.size(); i++)
          if (fr[i] != 0) cv[i] /= fr[i];
        auto last_at_least = [&](float th) {
          int i = sz - 1;
          for (; i >= 0; i--)
            if (cv[i] >= th) break;
          return i;
        };
        const int radius05 = last_at_least(0.5f), radius03 = last_at_least(0.3f);
        std::printf("the calculated radius of frc0.5 is %d\nthe calculated radius of frc0.3 is %d\n", radius05, radius03);
        if (radius05 > radius03 || radius03 == 0) {
          valid = false;
        } else {
          r05 = (float)radius05 / (float)sz;
          r03 = (float)radius03 / (float)sz;
        }
      }
      if (!valid) {
        std::printf("the FRC files are not valid\nAssign radius : %d\n", big_step_radius);
        radius_given = true;
        radius = big_step_radius;
      }
    }
    if (radius_given) {
      r05 = (float)((float)radius / ((float)Sx / 2.0));
      r03 = 0.f;
    }
    half.resize((size_t)Sx * (Sx / 2 + 1));
    int rc = icon_mask3_weights(Sx, thita.data(), (int)thita.size(), r05, r03, radius_given ? 1 : 0, gauss_width,
                                mean_step, half.data());
    if (rc) return rc;
  }
  // slices in batches of up to 64 on the device
  const int B = 64;
  std::vector<float> buf((size_t)B * Sx * Sy);
  float* d = nullptr;
  if (!no_mask && (hipSetDevice(device) != hipSuccess || hipMalloc(&d, sizeof(float) * buf.size()) != hipSuccess))
    return merr(ICON_ERR_DEVICE, "icon_mask3: device allocation failed");
  int rc = ICON_OK;
  std::string err;
  for (int z0 = 0; z0 < nslice && rc == ICON_OK; z0 += B) {
    const int cnt = std::min(B, nslice - z0);
    for (int q = 0; q < cnt && rc == ICON_OK; q++) {
      FILE* fp = std::fopen(mid(slice_begin + z0 + q).c_str(), "rb");
      MrcHeader h;
      if (!fp || !mrc::read_header(fp, &h) || h.nx != Sx || h.ny != Sy ||
          !mrc::read_z_section(fp, h, 0, buf.data() + (size_t)q * Sx * Sy, &err))
        rc = merr(ICON_ERR_IO, "Can not open infile " + mid(slice_begin + z0 + q) + "!");
      if (fp) std::fclose(fp);
    }
    if (rc) break;
    if (!no_mask) {
      if (hipMemcpy(d, buf.data(), sizeof(float) * cnt * Sx * Sy, hipMemcpyHostToDevice) != hipSuccess) {
        rc = merr(ICON_ERR_DEVICE, "icon_mask3: upload failed");
        break;
      }
      if ((rc = icon_mask3_filter(device, d, Sx, cnt, half.data(), nullptr))) break;
      if (hipMemcpy(buf.data(), d, sizeof(float) * cnt * Sx * Sy, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = merr(ICON_ERR_DEVICE, "icon_mask3: download failed");
        break;
      }
    }
    for (int q = 0; q < cnt; q++)
      std::memcpy(&out[(size_t)(z0 + q) * thickness * Sx], &buf[((size_t)q * Sy + ystart) * Sx],
                  sizeof(float) * thickness * Sx);
  }
  if (d) (void)hipFree(d);
  if (rc) return rc;
  // header of the first slice with nx, ny = thickness, nz = slices, float
  if (!mrc::write_float(out_path, head, out.data(), Sx, thickness, nslice, &err)) return merr(ICON_ERR_IO, err);
  return ICON_OK;
}

}  // extern "C"
