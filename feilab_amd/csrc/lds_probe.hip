// LDS bank probe for the 16-byte accesses of the U-Net convolution
// (ds_read_b128 / ds_write_b128): each wavefront accesses 16 bytes per lane
// at the addresses of one pattern, 4096 times, and the probe prints the LDS
// bytes per clock per CU (at 2.4 GHz) that the pattern sustains.  Lane l is
// row r = l % 16, piece g = l / 16 of a tile of rows, as the convolution's
// operand reads and staging writes are.  Build: make -C feilab_amd/csrc
// lds-probe; run: feilab_amd/lds-probe.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kIters = 4096;
constexpr int kPatterns = 9;
const char* const kNames[kPatterns] = {
    "consecutive 16 B per lane",
    "rows of 64 B",
    "rows of 80 B",
    "rows of 96 B",
    "rows of 144 B",
    "rows of 80 B, 4 lanes per row",
    "rows of 64 B, piece ^ ((row >> 2) & 2)",
    "rows of 64 B, piece ^ ((row >> 1) & 3)",
    "rows of 64 B, piece ^ ((row >> 2) & 3)",
};

__device__ int pattern(int p, int l) {
  const int r = l & 15, g = l >> 4;
  switch (p) {
    case 0: return l * 16;
    case 1: return r * 64 + g * 16;
    case 2: return r * 80 + g * 16;
    case 3: return r * 96 + g * 16;
    case 4: return r * 144 + g * 16;
    case 5: return (l >> 2) * 80 + (l & 3) * 16;
    case 6: return r * 64 + 16 * (g ^ ((r >> 2) & 2));
    case 7: return r * 64 + 16 * (g ^ ((r >> 1) & 3));
    default: return r * 64 + 16 * (g ^ ((r >> 2) & 3));
  }
}

// every address below is a multiple of 16 inside the 32 KB array
__global__ __launch_bounds__(256) void k_read(int p, unsigned* out) {
  __shared__ __align__(16) unsigned char lds[32768];
  const int t = threadIdx.x;
  for (int i = t; i < 32768 / 4; i += 256) reinterpret_cast<unsigned*>(lds)[i] = i * 2654435761u;
  __syncthreads();
  int a = pattern(p, t & 63) + (t >> 6) * 4096;
  uint4 acc = make_uint4(0, 0, 0, 0);
  for (int it = 0; it < kIters; it++) {
    asm volatile("" : "+v"(a));   // one LDS access per iteration
    const uint4 v = *reinterpret_cast<const uint4*>(lds + ((a + (it & 3) * 1024) & 32767));
    acc.x ^= v.x;
    acc.y += v.y;
    acc.z ^= v.z;
    acc.w += v.w;
  }
  out[blockIdx.x * 256 + t] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

__global__ __launch_bounds__(256) void k_write(int p, unsigned* out) {
  __shared__ __align__(16) unsigned char lds[32768];
  const int t = threadIdx.x;
  int a = pattern(p, t & 63) + (t >> 6) * 4096;
  uint4 v = make_uint4(t, t + 1, t + 2, t + 3);
  for (int it = 0; it < kIters; it++) {
    asm volatile("" : "+v"(a));
    *reinterpret_cast<uint4*>(lds + ((a + (it & 3) * 1024) & 32767)) = v;
    v.x += 1;
  }
  __syncthreads();
  out[blockIdx.x * 256 + t] = reinterpret_cast<unsigned*>(lds)[t];
}

}  // namespace

int main() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev))
    return 1;
  const int blocks = 4 * cus;
  unsigned* d = nullptr;
  if (hipMalloc(&d, sizeof(unsigned) * blocks * 256) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  printf("| pattern | ds_read_b128 B/clk/CU | ds_write_b128 B/clk/CU |\n|---|---:|---:|\n");
  for (int p = 0; p < kPatterns; p++) {
    double rate[2];
    for (int w = 0; w < 2; w++) {
      auto launch = [&] {
        if (w == 0)
          hipLaunchKernelGGL(k_read, dim3(blocks), dim3(256), 0, 0, p, d);
        else
          hipLaunchKernelGGL(k_write, dim3(blocks), dim3(256), 0, 0, p, d);
      };
      launch();
      (void)hipEventRecord(e0);
      for (int rep = 0; rep < 5; rep++) launch();
      (void)hipEventRecord(e1);
      if (hipEventSynchronize(e1) != hipSuccess) return 2;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double bytes = 5.0 * blocks * 256.0 * 16.0 * kIters;
      rate[w] = bytes / (ms * 1e-3) / cus / 2.4e9;
    }
    printf("| %s | %.0f | %.0f |\n", kNames[p], rate[0], rate[1]);
  }
  (void)hipFree(d);
  return 0;
}
