// Internal plan state of the MI355X ICON core (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/icon_mi355x.h"

namespace icon {

constexpr int kTaps = 14;      // Gaussian window: 2m+2, m = 6 (nufft_gpu_v8.cuh:5)
constexpr int kWindowM = 6;
constexpr int kTapsKB = 6;     // Kaiser-Bessel window of width 6
constexpr int kMaxSlot = 3;    // node y-taps split over at most 3 column tiles (tile >= 8)

// Host-side geometry of one tilt series, computed exactly as the reference's
// setup does (float/double placement included).
struct Geometry {
  int N0 = 0, N1 = 0, nz = 0, M = 0, n0 = 0, n1 = 0;
  int window = ICON_WINDOW_GAUSS;  // ICON_WINDOW_GAUSS or ICON_WINDOW_KB6
  int K = kTaps;                  // taps per dimension: 14 (Gaussian) or 6 (Kaiser-Bessel)
  // window shape per dimension: Gaussian b and bf, or Kaiser-Bessel beta (b)
  float sig0 = 0, sig1 = 0, b0 = 0, b1 = 0, bf0 = 0, bf1 = 0;
  int iomit = 0;                  // tilt whose nodes give the reprojection
  float omit = 0;
  int cv_skip = -1;               // cross-validation plan: input row left out of the data, else -1
  std::vector<float> theta;       // negated angles (MainCode.cu:42)
  std::vector<float> nodes;       // 2*M, (x, y) per node j = i*N0 + k
  std::vector<float> w;           // M density weights
  std::vector<float> ck;          // N0*N1 deapodisation, index X*N1+Y
  std::vector<float> ck1;         // N0
  std::vector<float> phix, phiy;  // M*K window taps
  std::vector<int> ix0, iy0;      // M: wrapped first tap index
  std::vector<float> phi1;        // N0*K (1D plan)
  std::vector<int> i10;           // N0
};

// Column-tile bin lists for the spread / interpolation.
struct Bins {
  int C = 16;                     // grid columns per tile
  int nblk = 0;                   // n1 / C
  std::vector<int> off;           // nblk+1
  // (node j, t0 | t1<<8 | slot<<16 | col0<<24, first x-tap x0, 0); a tile's
  // entries are sorted by x0 (then node)
  std::vector<int4> ent;
  std::vector<int> xoff;          // nblk*(n0+1): first entry of the tile with x0 >= x
  std::vector<int8_t> nslot;      // M
  std::vector<float> px;          // E*14: x-taps of the entry's node (interleaved per tile)
  std::vector<float> py;          // E*C: y-tap of each tile column, 0 outside the run (interleaved per tile)
  std::vector<int> of_node;       // M*3: entries of node j by slot (-1: none)
  // The same entries and records in the interpolation's order: within a tile,
  // groups of 16 entries whose first x-taps differ mod 16 where possible
  // (records interleaved by field per tile as above).
  std::vector<int4> fent;
  std::vector<float> fpx, fpy;
  // Launch order of the spread's tiles: by entry count, largest first (ties
  // by index).  A spread block costs about its tile's entries, and the few
  // dense tiles around ky = 0 (up to 10x the median block) would otherwise
  // run at the start and the very end of the tile-major order, the last of
  // them alone on a nearly idle device.
  std::vector<int> order;
};

// The grid columns [zlo, zhi) (FFT order, a multiple of C at both ends) of
// the run of column tiles without bin entries around column n1/2.  Every
// tilt line's nodes reach ky in [-k, k], so the tiles the nodes reach form a
// band around column 0 and the free ones a run around n1/2.  Spread grids
// are zero there and no interpolation reads a forward grid there, so the
// row kernels neither load nor store those columns and the spread kernel
// does not write them.  Empty (0, 0) when the tile holding column n1/2 is
// reached.
inline void zero_band(const Bins& b, int* zlo, int* zhi) {
  *zlo = *zhi = 0;
  const int mid = b.nblk / 2;
  auto empty = [&](int t) { return b.off[t] == b.off[t + 1]; };
  if (b.nblk < 2 || !empty(mid)) return;
  int lo = mid, hi = mid + 1;
  while (lo > 0 && empty(lo - 1)) lo--;
  while (hi < b.nblk && empty(hi)) hi++;
  *zlo = lo * b.C;
  *zhi = hi * b.C;
}

// Entry count from which a spread tile splits each row pair's entries
// between two threads: twice the median count of the tiles that have
// entries, so that only the dense tiles around ky = 0 reach it (22 of the
// 114 tiles with entries at 512^2, 41 tilts +-60 deg; 1.5x the median, 38
// tiles, and 3x, 8 tiles, measured slower).
inline int split_min_entries(const Bins& b) {
  std::vector<int> n;
  for (int t = 0; t < b.nblk; t++)
    if (b.off[t + 1] > b.off[t]) n.push_back(b.off[t + 1] - b.off[t]);
  if (n.empty()) return 0x7fffffff;
  std::nth_element(n.begin(), n.begin() + n.size() / 2, n.end());
  return std::max(1, 2 * n[n.size() / 2]);
}

// Fills g; returns false (with a message) on an invalid geometry.
bool build_geometry(Geometry& g, int N0, int N1, int nz, const float* tilt_deg, std::string* err, bool cv = false,
                    int window = ICON_WINDOW_GAUSS);
void build_bins(const Geometry& g, int C, Bins& b);
void grid_size(int N, float* sigma, int* n);
// Sets the message icon_last_error() returns on the calling thread.
void set_last_error(const char* msg);

struct GraphKey {
  int S = -1;
  const void* sino = nullptr;
  void* m = nullptr;
  const void* m_init = nullptr;
  void* reproj = nullptr;
  void* resid = nullptr;
  void* alpha = nullptr;
  int pre = 0, icon = 0, post = 0, data_type = 0, reuse = 0;
  float thr = 0;
  bool operator==(const GraphKey& o) const {
    return S == o.S && sino == o.sino && m == o.m && m_init == o.m_init && reproj == o.reproj &&
           resid == o.resid && alpha == o.alpha && pre == o.pre && icon == o.icon &&
           post == o.post && data_type == o.data_type && reuse == o.reuse && thr == o.thr;
  }
};

// A few captured solves per plan (a warm-started solve alternates between two
// or three schedules); least recently used is replaced.
constexpr int kGraphCache = 4;

}  // namespace icon

// Opaque plan of the C ABI.
struct icon_plan {
  int device = 0;
  icon::Geometry g;
  icon::Bins bins;
  int S_max = 0;
  hipStream_t stream = nullptr;     // plan-owned stream (capturable)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // shared tables
  float* d_ck = nullptr;
  float* d_ck1 = nullptr;
  float* d_w = nullptr;
  float* d_phix = nullptr;
  float* d_phiy = nullptr;
  int* d_ix0 = nullptr;
  int* d_ent_off = nullptr;
  int* d_tile_order = nullptr;      // Bins::order
  int4* d_ent = nullptr;
  int* d_xoff = nullptr;
  float* d_ent_px = nullptr;
  float* d_ent_py = nullptr;
  int4* d_fent = nullptr;
  float* d_fent_px = nullptr;
  float* d_fent_py = nullptr;
  int* d_ent_of_node = nullptr;
  int8_t* d_nslot = nullptr;
  float* d_phi1 = nullptr;
  int* d_i10 = nullptr;
  const float2* d_tw0 = nullptr;    // twiddles for n0 (shared cache)
  const float2* d_tw1 = nullptr;    // twiddles for n1
  // per-slice work buffers (capacity S_max)
  float2* d_b = nullptr;            // [S][M] Fourier-slice data
  float2* d_fw = nullptr;           // [S][E] weighted node value of every bin entry
  float2* d_part = nullptr;         // [S][M][kMaxSlot] interpolation partials
  float2* d_ahwb = nullptr;         // [S][N0][N1]
  float* d_grad = nullptr;          // [S][N0][N1] Re(rhat), the gradient the update takes
  float2* d_tmp = nullptr;          // [S][N0][n1] row-transform intermediate
  float2* d_frep = nullptr;         // [S][N0] forward values at the omit tilt
  double* d_rowpart = nullptr;      // [S][N0] per-row partial sums
  double* d_bnorm = nullptr;        // [S]
  double* d_alpha = nullptr;        // [S]
  int64_t device_bytes = 0;
  // captured solves
  hipGraphExec_t graph_exec[icon::kGraphCache] = {};
  icon::GraphKey graph_key[icon::kGraphCache];
  int64_t graph_used[icon::kGraphCache] = {};
  int64_t graph_clock = 0;
  // sinograms whose b and A^H W b the per-slice buffers hold (reuse_data)
  const void* data_sino = nullptr;
  int data_S = 0;
};
