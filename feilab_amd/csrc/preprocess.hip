// Tilt-series pre-processing on the device (ICONPreProcess,
// sourceCode/ICONPreProcess/ICONPreProcess.c:31-86,256-302): per projection
// image, subtract the histogram mode of its central region and, given the
// tilt angles and the specimen thickness, scale it to variance
// 0.33 thickness / cos(theta).  This fixes the input scale that ICON's
// threshold (-thr) assumes.
//
// One 1024-thread block per image: FP64 block reductions for min/max, mean
// and variance, a 256-bin histogram in LDS (integer LDS atomics), and the
// pointwise updates in the reference's double-then-float order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icon_mi355x.h"
#include "mrc_io.h"

namespace {

constexpr int kThreads = 1024;
constexpr int kBins = 256;

template <typename Op>
__device__ double block_reduce(double v, double* red, Op op) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = op(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kThreads) void k_preprocess(float* stack, int nx, int ny, const float* tilt_deg,
                                                         int thickness, double* modes) {
  __shared__ double red[kThreads];
  __shared__ int hist[kBins];
  const int k = blockIdx.x;
  float* in = stack + (size_t)k * nx * ny;
  const long size = (long)nx * ny;
  const int xe = (int)(nx * 0.1), ye = (int)(ny * 0.1);
  const int cw = nx - 2 * xe;
  const long csize = (long)cw * (ny - 2 * ye);
  // min / max of the central region
  double mn = in[(size_t)ye * nx + xe], mx = mn;
  for (long q = threadIdx.x; q < csize; q += kThreads) {
    const double v = in[(size_t)(ye + q / cw) * nx + xe + q % cw];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  mn = block_reduce(mn, red, [](double a, double b) { return a < b ? a : b; });
  mx = block_reduce(mx, red, [](double a, double b) { return a > b ? a : b; });
  const double step = (mx - mn) / (double)kBins;
  double mode = mn;
  if (step > 0) {
    for (int b = threadIdx.x; b < kBins; b += kThreads) hist[b] = 0;
    __syncthreads();
    for (long q = threadIdx.x; q < csize; q += kThreads) {
      const double v = in[(size_t)(ye + q / cw) * nx + xe + q % cw];
      const int b = (int)floor((v - mn) / step);
      if (b >= 0 && b < kBins) atomicAdd(&hist[b], 1);   // the maximum (b == 256) is not counted
    }
    __syncthreads();
    // first bin of the largest count
    int bi = 0, best = -1000;
    for (int b = 0; b < kBins; b++)
      if (best < hist[b]) { best = hist[b]; bi = b; }
    mode = bi * step + mn;
  }
  if (threadIdx.x == 0 && modes) modes[k] = mode;
  for (long i = threadIdx.x; i < size; i += kThreads) in[i] = (float)((double)in[i] - mode);
  if (!tilt_deg || thickness <= 0) return;
  __syncthreads();
  double s = 0;
  const float fsize = (float)size;
  for (long i = threadIdx.x; i < size; i += kThreads) s += (double)(in[i] / fsize);   // float division, as the reference
  const double mean = block_reduce(s, red, [](double a, double b) { return a + b; });
  double v2 = 0;
  for (long i = threadIdx.x; i < size; i += kThreads) {
    const double d = (double)in[i] - mean;
    v2 += d * d / (double)size;
  }
  const double var = block_reduce(v2, red, [](double a, double b) { return a + b; });
  const double angle = -(double)tilt_deg[k];
  const double factor = 0.33 * thickness / cos(angle / 180.0 * 3.1415926);
  const double sv = sqrt(var), sf = sqrt(factor);
  for (long i = threadIdx.x; i < size; i += kThreads) in[i] = (float)((double)in[i] / sv * sf);
}

thread_local std::string g_perr;

int perr(int code, const std::string& m) {
  g_perr = m;
  return code;
}

// Device memory of icon_preprocess_noalloc, in place of the two buffers that
// icon_preprocess allocates per call: the subtracted modes (double, staged
// out) at offset 0, the tilt angles (float, staged in) after them.  Each
// device holds its own copy until the process ends.
constexpr int kNoAllocMaxImages = 4096;
constexpr size_t kScratchModes = 0;
constexpr size_t kScratchTilts = sizeof(double) * kNoAllocMaxImages;
constexpr size_t kScratchBytes = kScratchTilts + sizeof(float) * kNoAllocMaxImages;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on each device, resolved once per device with
// that device current (the array's name in host code is only a stand-in).
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;

hipError_t scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" {

const char* icon_preprocess_error(void) { return g_perr.c_str(); }

int icon_preprocess(int device, float* d_stack, int nx, int ny, int nz, const float* tilt_deg, int thickness,
                    double* modes, void* stream) {
  if (!d_stack || nx < 1 || ny < 1 || nz < 1) return perr(ICON_ERR_ARG, "icon_preprocess: bad argument");
  if (hipSetDevice(device) != hipSuccess) return perr(ICON_ERR_ARG, "icon_preprocess: no such device");
  hipStream_t st = (hipStream_t)stream;
  float* d_tilt = nullptr;
  double* d_modes = nullptr;
  const bool norm = tilt_deg && thickness > 0;
  hipError_t e = hipSuccess;
  if (norm) e = hipMalloc(&d_tilt, sizeof(float) * nz);
  if (e == hipSuccess && norm) e = hipMemcpy(d_tilt, tilt_deg, sizeof(float) * nz, hipMemcpyHostToDevice);
  if (e == hipSuccess && modes) e = hipMalloc(&d_modes, sizeof(double) * nz);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_preprocess, dim3(nz), dim3(kThreads), 0, st, d_stack, nx, ny, d_tilt, norm ? thickness : 0,
                       d_modes);
    e = hipGetLastError();
  }
  if (e == hipSuccess && modes) e = hipMemcpyAsync(modes, d_modes, sizeof(double) * nz, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d_tilt) (void)hipFree(d_tilt);
  if (d_modes) (void)hipFree(d_modes);
  if (e != hipSuccess) return perr(ICON_ERR_DEVICE, std::string("icon_preprocess: ") + hipGetErrorString(e));
  return ICON_OK;
}

int icon_preprocess_noalloc(int device, float* d_stack, int nx, int ny, int nz, const float* tilt_deg, int thickness,
                            double* modes, void* stream) {
  if (!d_stack || nx < 1 || ny < 1 || nz < 1) return perr(ICON_ERR_ARG, "icon_preprocess_noalloc: bad argument");
  if (nz > kNoAllocMaxImages)
    return perr(ICON_ERR_CAPACITY, "icon_preprocess_noalloc: " + std::to_string(nz) + " images exceed the " +
                                       std::to_string(kNoAllocMaxImages) + " its device scratch holds per call");
  if (hipSetDevice(device) != hipSuccess) return perr(ICON_ERR_ARG, "icon_preprocess_noalloc: no such device");
  unsigned char* base = nullptr;
  hipError_t e = scratch_on_current_device(&base);
  if (e != hipSuccess)
    return perr(ICON_ERR_DEVICE, std::string("icon_preprocess_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
  hipStream_t st = (hipStream_t)stream;
  const bool norm = tilt_deg && thickness > 0;
  // every launch writes all nz modes and reads only the nz tilts staged here,
  // so nothing left in the scratch by an earlier call is read
  float* d_tilt = norm ? reinterpret_cast<float*>(base + kScratchTilts) : nullptr;
  double* d_modes = modes ? reinterpret_cast<double*>(base + kScratchModes) : nullptr;
  if (norm) e = hipMemcpyAsync(d_tilt, tilt_deg, sizeof(float) * nz, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_preprocess, dim3(nz), dim3(kThreads), 0, st, d_stack, nx, ny, d_tilt, norm ? thickness : 0,
                       d_modes);
    e = hipGetLastError();
  }
  if (e == hipSuccess && modes) e = hipMemcpyAsync(modes, d_modes, sizeof(double) * nz, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return perr(ICON_ERR_DEVICE, std::string("icon_preprocess_noalloc: ") + hipGetErrorString(e));
  return ICON_OK;
}

int icon_preprocess_file(int device, const char* in_path, const char* out_path, const char* tilt_path,
                         int thickness) {
  if (!in_path || !out_path) return perr(ICON_ERR_ARG, "icon_preprocess_file: input or output missing");
  FILE* fp = std::fopen(in_path, "rb");
  if (!fp) return perr(ICON_ERR_IO, std::string("Cannot open file '") + in_path + "'");
  MrcHeader h;
  std::string err;
  if (!mrc::read_header(fp, &h)) {
    std::fclose(fp);
    return perr(ICON_ERR_IO, "cannot read the MRC header");
  }
  const size_t sec = (size_t)h.nx * h.ny;
  std::vector<float> stack(sec * h.nz);
  for (int k = 0; k < h.nz; k++)
    if (!mrc::read_z_section(fp, h, k, stack.data() + k * sec, &err)) {
      std::fclose(fp);
      return perr(ICON_ERR_IO, err);
    }
  std::fclose(fp);
  std::vector<float> tilts;
  if (tilt_path && thickness > 0) {
    tilts.assign(1000, 0.f);
    int n = 0;
    if (icon_read_tilts(tilt_path, tilts.data(), (int)tilts.size(), &n) != ICON_OK)
      return perr(ICON_ERR_IO, std::string("Cannot open file '") + tilt_path + "'");
    if (n < h.nz) return perr(ICON_ERR_ARG, "tilt file has fewer angles than the stack has sections");
  }
  float* d = nullptr;
  if (hipSetDevice(device) != hipSuccess || hipMalloc(&d, sizeof(float) * stack.size()) != hipSuccess)
    return perr(ICON_ERR_DEVICE, "icon_preprocess_file: device allocation failed");
  int rc = ICON_OK;
  if (hipMemcpy(d, stack.data(), sizeof(float) * stack.size(), hipMemcpyHostToDevice) != hipSuccess) {
    rc = perr(ICON_ERR_DEVICE, "icon_preprocess_file: upload failed");
  } else {
    rc = icon_preprocess(device, d, h.nx, h.ny, h.nz, tilts.empty() ? nullptr : tilts.data(), thickness, nullptr,
                         nullptr);
    if (rc == ICON_OK && hipMemcpy(stack.data(), d, sizeof(float) * stack.size(), hipMemcpyDeviceToHost) != hipSuccess)
      rc = perr(ICON_ERR_DEVICE, "icon_preprocess_file: download failed");
  }
  (void)hipFree(d);
  if (rc) return rc;
  // header copied from the input, mode float, statistics recomputed
  if (!mrc::write_float(out_path, h, stack.data(), h.nx, h.ny, h.nz, &err)) return perr(ICON_ERR_IO, err);
  return ICON_OK;
}

}  // extern "C"
