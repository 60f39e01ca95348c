// Host-side geometry of an ICON plan: oversampling, Gaussian window tables,
// deapodisation, radial nodes, density weights and the column-tile bin lists
// used by the spread/interpolation kernels.
//
// The float/double placement follows the reference setup so that the device
// tables equal the reference's (paths relative to
// tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
//   oversampling          MainCode.cu:63-71
//   window b, bf          nufft_gpu_v8.cu:471-476, 1330-1331
//   window taps           nufft_gpu_v8.cu:221-235 (2D), 1242-1256 (1D)
//   deapodisation         nufft_gpu_v8.cu:22-42 (2D), 1258-1271 (1D)
//   radial nodes          px2D.cu:8-47 (pi = 3.14159265, px2D.cuh:8)
//   1D nodes              MainCode.cu:75-78
//   density weights       weight.cu:8-41
//   omitted angle         MainCode.cu:46-49
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "icon_plan.h"

namespace icon {

namespace {
constexpr double kPI = 3.1415926535898;   // nufft_gpu_v8.cuh:8
constexpr double kPiNodes = 3.14159265;   // px2D.cuh:8

// b = 2 sigma m / ((2 sigma - 1) pi): float numerator, double denominator,
// stored in float; bf = 1/sqrt(pi b) stored in float.
void window_shape(float sigma, int m, float* b, float* bf) {
  const float num = 2.0f * sigma * (float)m;
  const double den = (double)(2.0f * sigma - 1.0f) * kPI;
  *b = (float)((double)num / den);
  *bf = (float)(1.0 / std::sqrt(kPI * (double)*b));
}

// One factor of the deapodisation exp(-b (pi k / n)^2): double exponent with
// the float b and float n, stored in float.
float deapod(int k, int N, float sigma, int m) {
  const float n = std::ceil((float)N * sigma);
  float b, bf;
  window_shape(sigma, m, &b, &bf);
  const double a = kPI * (double)k / (double)n;
  return (float)std::exp(-(double)b * (a * a));
}

// Window taps of one coordinate (float arithmetic, expf).
void taps(float c, int N, float sigma, float b, float bf, int m, float* phi, int* first) {
  const int n = (int)std::ceil((float)N * sigma);
  const float fl = std::floor((float)n * c);
  for (int t = 0; t < 2 * m + 2; t++) {
    const int l = (int)(fl - (float)m + (float)t);
    const float x1 = c - (float)l / (float)n;
    const float nx1 = (float)n * x1;
    phi[t] = bf * std::exp(-(nx1 * nx1 / b));
    if (t == 0) *first = (l + n) % n;
  }
}
// Kaiser-Bessel window of width W = 6 grid points (6 taps per dimension),
// an alternative to the reference's Gaussian: phi(d) = I0(beta sqrt(1 -
// (2d/W)^2)) / Z for |d| <= W/2, with beta = pi sqrt((W/sigma)^2 (sigma -
// 1/2)^2 - 0.8) (Beatty, Nishimura and Pauly, IEEE TMI 24(6), 2005) and
// Z = W sinh(beta) / beta, so that the deapodisation (its Fourier
// transform W sinh(a) / a, a = sqrt(beta^2 - (pi W k / n)^2), over Z) is 1
// at k = 0 like the Gaussian's.  Computed in double, stored in float.
double bessel_i0(double x) {
  // power series sum_k ((x/2)^k / k!)^2; |x| <= 20 here
  double term = 1.0, sum = 1.0;
  const double q = 0.25 * x * x;
  for (int k = 1; k < 200 && term > 1e-17 * sum; k++) {
    term *= q / ((double)k * (double)k);
    sum += term;
  }
  return sum;
}

float kb_beta(float sigma) {
  const double W = kTapsKB, s = sigma;
  return (float)(kPI * std::sqrt((W / s) * (W / s) * (s - 0.5) * (s - 0.5) - 0.8));
}

double kb_norm(double beta) { return kTapsKB * std::sinh(beta) / beta; }

float deapod_kb(int k, int N, float sigma, float beta) {
  const double n = std::ceil((float)N * sigma);
  const double W = kTapsKB, bt = beta;
  const double u = kPI * W * (double)k / n;
  const double a2 = bt * bt - u * u;
  double v;
  if (a2 > 0) v = W * std::sinh(std::sqrt(a2)) / std::sqrt(a2);
  else if (a2 < 0) v = W * std::sin(std::sqrt(-a2)) / std::sqrt(-a2);
  else v = W;
  return (float)(v / kb_norm(bt));
}

// Kaiser-Bessel taps of one coordinate: grid points floor(n c) - 2 .. + 3.
void taps_kb(float c, int N, float sigma, float beta, float* phi, int* first) {
  const int n = (int)std::ceil((float)N * sigma);
  const double u = (double)n * (double)c;
  const int l0 = (int)std::floor(u) - (kTapsKB / 2 - 1);
  const double Z = kb_norm(beta);
  for (int t = 0; t < kTapsKB; t++) {
    const double d = u - (double)(l0 + t);
    const double r = 1.0 - (2.0 * d / kTapsKB) * (2.0 * d / kTapsKB);
    phi[t] = r > 0 ? (float)(bessel_i0((double)beta * std::sqrt(r)) / Z) : 0.f;
  }
  *first = ((l0 % n) + n) % n;
}
}  // namespace

void grid_size(int N, float* sigma, int* n) {
  const float uptop = (float)std::pow(2.0, std::ceil(std::log10((double)N) / std::log10(2.0)));
  const float s = (float)(2.0 * uptop / (double)N);
  *sigma = s;
  *n = (int)std::ceil(s * (float)N);
}

bool build_geometry(Geometry& g, int N0, int N1, int nz, const float* tilt_deg, std::string* err, bool cv,
                    int window) {
  if (window != ICON_WINDOW_GAUSS && window != ICON_WINDOW_KB6) {
    if (err) *err = "icon_plan_create: unknown window";
    return false;
  }
  if (N0 < 8 || N1 < 8 || nz < (cv ? 3 : 2)) {
    if (err) *err = cv ? "icon_plan_create_cv: need nx >= 8, thickness >= 8 and at least 3 tilts"
                       : "icon_plan_create: need nx >= 8, thickness >= 8 and at least 2 tilts";
    return false;
  }
  if (nz > 1000) {
    if (err) *err = "icon_plan_create: more than 1000 tilts";
    return false;
  }
  g.N0 = N0; g.N1 = N1; g.nz = nz; g.M = N0 * nz;
  grid_size(N0, &g.sig0, &g.n0);
  grid_size(N1, &g.sig1, &g.n1);
  for (int n : {g.n0, g.n1}) {
    if (n < 128 || n > 2048 || (n & (n - 1)) != 0) {
      if (err) {
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "icon_plan_create: oversampled grid %d unsupported (widths and thickness 33..1024 "
                      "give grids 128..2048)", n);
        *err = buf;
      }
      return false;
    }
  }
  g.window = window;
  const bool kb = window == ICON_WINDOW_KB6;
  g.K = kb ? kTapsKB : kTaps;
  if (kb) {
    g.b0 = kb_beta(g.sig0);
    g.b1 = kb_beta(g.sig1);
    g.bf0 = g.bf1 = 0.f;
  } else {
    window_shape(g.sig0, kWindowM, &g.b0, &g.bf0);
    window_shape(g.sig1, kWindowM, &g.b1, &g.bf1);
  }

  // angles are negated on read (MainCode.cu:42); the omitted tilt is the
  // first of least |theta| (MainCode.cu:46-49)
  g.theta.resize(nz);
  for (int i = 0; i < nz; i++) g.theta[i] = -tilt_deg[i];
  float om = g.theta[0];
  for (int i = 0; i < nz; i++)
    if (std::fabs(om) > std::fabs(g.theta[i])) om = g.theta[i];
  g.omit = om;
  g.iomit = 0;
  for (int i = 0; i < nz; i++)
    if (g.theta[i] == om) { g.iomit = i; break; }
  // cross-validation plan (MainCode_crossV.cu:45-60): the omitted tilt leaves
  // the data; the remaining nz-1 tilts keep their order and their own weights,
  // and the omitted angle is appended as a reprojection line of weight 0
  // (its nodes then contribute exactly nothing to A^H W, R or alpha)
  g.cv_skip = -1;
  int nw = nz;   // tilts that carry data
  if (cv) {
    g.cv_skip = g.iomit;
    for (int i = g.iomit; i < nz - 1; i++) g.theta[i] = g.theta[i + 1];
    g.theta[nz - 1] = om;
    g.iomit = nz - 1;
    nw = nz - 1;
  }

  // radial nodes: float accumulation of the radius, double trig
  const int M = g.M;
  g.nodes.resize(2 * (size_t)M);
  {
    std::vector<float> f(N0);
    const float step = (float)(1.0 / N0);
    f[0] = -0.5f;
    for (int k = 1; k < N0; k++) f[k] = f[k - 1] + step;
    for (int i = 0; i < nz; i++) {
      const double a = (double)g.theta[i] * kPiNodes / 180.0;
      const double c = std::cos(a), s = std::sin(a);
      for (int k = 0; k < N0; k++) {
        g.nodes[2 * ((size_t)i * N0 + k)] = (float)((double)f[k] * c);
        g.nodes[2 * ((size_t)i * N0 + k) + 1] = (float)((double)f[k] * s);
      }
    }
  }

  // density weights over the nw data tilts: |ramp| x angular spacing, centre
  // 180/(4 nw); a cross-validation plan's reprojection line has weight 0
  g.w.assign(M, 0.f);
  {
    const int center = N0 / 2;
    const float* th = g.theta.data();
    for (int i = 0; i < nw; i++) {
      const float span = (i == 0) ? (th[1] - th[0]) : (i == nw - 1) ? (th[nw - 2] - th[nw - 1])
                                                                     : (th[i - 1] - th[i + 1]);
      const double mult = (i == 0 || i == nw - 1) ? 2.0 : 1.0;
      for (int k = 0; k < N0; k++) {
        float v;
        if (k == center) v = (float)(180.0 / (4 * nw));
        else v = (float)(mult * std::fabs((double)(span * (float)(k - center))));
        g.w[(size_t)i * N0 + k] = v;
      }
    }
  }

  // 2D window taps and first indices
  const int K = g.K;
  auto node_taps = [&](float c, int N, float sigma, float b, float bf, float* phi, int* first) {
    if (kb) taps_kb(c, N, sigma, b, phi, first);
    else taps(c, N, sigma, b, bf, kWindowM, phi, first);
  };
  auto deapod_of = [&](int k, int N, float sigma, float b) {
    return kb ? deapod_kb(k, N, sigma, b) : deapod(k, N, sigma, kWindowM);
  };
  g.phix.resize((size_t)M * K);
  g.phiy.resize((size_t)M * K);
  g.ix0.resize(M);
  g.iy0.resize(M);
  for (int j = 0; j < M; j++) {
    node_taps(g.nodes[2 * (size_t)j], N0, g.sig0, g.b0, g.bf0, &g.phix[(size_t)j * K], &g.ix0[j]);
    node_taps(g.nodes[2 * (size_t)j + 1], N1, g.sig1, g.b1, g.bf1, &g.phiy[(size_t)j * K], &g.iy0[j]);
  }

  // deapodisation
  g.ck.resize((size_t)N0 * N1);
  {
    std::vector<float> cy(N1);
    for (int Y = 0; Y < N1; Y++) cy[Y] = deapod_of(Y - N1 / 2, N1, g.sig1, g.b1);
    for (int X = 0; X < N0; X++) {
      const float cx = deapod_of(X - N0 / 2, N0, g.sig0, g.b0);
      for (int Y = 0; Y < N1; Y++) g.ck[(size_t)X * N1 + Y] = cx * cy[Y];
    }
  }

  // 1D plan of the projection rows: nodes (k - N0/2)/N0, same sigma as dim 0
  g.ck1.resize(N0);
  g.phi1.resize((size_t)N0 * K);
  g.i10.resize(N0);
  for (int k = 0; k < N0; k++) {
    const float x = (float)(((double)(float)k - (double)(float)N0 / 2.0) / (double)(float)N0);
    node_taps(x, N0, g.sig0, g.b0, g.bf0, &g.phi1[(size_t)k * K], &g.i10[k]);
    g.ck1[k] = deapod_of(k - N0 / 2, N0, g.sig0, g.b0);
  }
  return true;
}

// Bins every node's run of y-taps by grid-column tile.  A node's K taps are
// consecutive columns (mod n1), so they fall into at most ceil((K-1)/C)+1 tiles;
// each (tile, node) pair is one entry recording the tap range, the node's
// partial-sum slot, the first column inside the tile and the node's first
// x-tap.  A tile's entries are sorted by first x-tap (then node), and xoff
// gives, per tile and grid row x, the first entry whose first x-tap is >= x,
// so the entries covering row x are those with first x-tap in [x-K+1, x].
void build_bins(const Geometry& g, int C, Bins& b) {
  const int K = g.K;
  b.C = C;
  b.nblk = g.n1 / C;
  b.off.assign(b.nblk + 1, 0);
  b.nslot.assign(g.M, 0);
  const int mask = g.n1 - 1;
  for (int j = 0; j < g.M; j++) {
    int prev = -1;
    for (int t = 0; t < K; t++) {
      const int blk = ((g.iy0[j] + t) & mask) / C;
      if (blk != prev) { b.off[blk + 1]++; prev = blk; }
    }
  }
  for (int i = 0; i < b.nblk; i++) b.off[i + 1] += b.off[i];
  b.order.resize(b.nblk);
  for (int i = 0; i < b.nblk; i++) b.order[i] = i;
  std::stable_sort(b.order.begin(), b.order.end(), [&](int u, int v) {
    return b.off[u + 1] - b.off[u] > b.off[v + 1] - b.off[v];
  });
  b.ent.resize(b.off[b.nblk]);
  std::vector<int> fill(b.off.begin(), b.off.end() - 1);
  for (int j = 0; j < g.M; j++) {
    int t = 0, slot = 0;
    while (t < K) {
      const int col = (g.iy0[j] + t) & mask;
      const int blk = col / C;
      int t1 = t;
      while (t1 + 1 < K && (((g.iy0[j] + t1 + 1) & mask) / C) == blk) t1++;
      const int col0 = col - blk * C;
      b.ent[fill[blk]++] = make_int4(j, t | (t1 << 8) | (slot << 16) | (col0 << 24), g.ix0[j], 0);
      slot++;
      t = t1 + 1;
    }
    b.nslot[j] = (int8_t)slot;
  }
  const int n0 = g.n0;
  b.xoff.assign((size_t)b.nblk * (n0 + 1), 0);
  for (int blk = 0; blk < b.nblk; blk++) {
    std::stable_sort(b.ent.begin() + b.off[blk], b.ent.begin() + b.off[blk + 1],
                     [](const int4& u, const int4& v) { return u.z < v.z; });
    int* xo = &b.xoff[(size_t)blk * (n0 + 1)];
    int e = b.off[blk];
    for (int x = 0; x <= n0; x++) {
      while (e < b.off[blk + 1] && b.ent[e].z < x) e++;
      xo[x] = e;
    }
  }
  // self-contained records: the node's x-taps, the y-tap under every column
  // of the tile (zero outside the run) and the node -> entries map.  Within a
  // tile the records are interleaved by field (x-tap pairs, y-tap quads) so
  // that consecutive entries of one field are contiguous.
  const size_t E = b.ent.size();
  b.px.assign(E * K, 0.f);
  b.py.assign(E * (size_t)C, 0.f);
  b.of_node.assign((size_t)g.M * kMaxSlot, -1);
  for (int blk = 0; blk < b.nblk; blk++) {
    const size_t e0 = b.off[blk], ne = b.off[blk + 1] - e0;
    for (size_t e = e0; e < e0 + ne; e++) {
      const int4 r = b.ent[e];
      const int j = r.x, t0 = r.y & 0xff, t1 = (r.y >> 8) & 0xff, slot = (r.y >> 16) & 0xff;
      const int col0 = (r.y >> 24) & 0xff;
      for (int t = 0; t < K; t++)
        b.px[e0 * K + ((t / 2) * ne + (e - e0)) * 2 + (t & 1)] = g.phix[(size_t)j * K + t];
      for (int t = t0; t <= t1; t++) {
        const int c = col0 + t - t0;
        b.py[e0 * C + ((c / 4) * ne + (e - e0)) * 4 + (c & 3)] = g.phiy[(size_t)j * K + t];
      }
      b.of_node[(size_t)j * kMaxSlot + slot] = (int)e;
    }
  }
  // Interpolation order.  The interpolation reads rows of 80 bytes from the
  // LDS tile, rows x0 .. x0+K-1 of entry e in lane e; two lanes of a 16-lane
  // group whose first taps differ by a multiple of 16 rows hit the same
  // banks.  Each group of 16 takes the next entry (by x0) of every residue
  // x0 mod 16 in turn, so its first taps differ mod 16 while entries remain
  // in 16 residues, and stay close to each other.
  b.fent.resize(E);
  b.fpx.assign(E * K, 0.f);
  b.fpy.assign(E * (size_t)C, 0.f);
  for (int blk = 0; blk < b.nblk; blk++) {
    const size_t e0 = b.off[blk], ne = b.off[blk + 1] - e0;
    std::vector<size_t> bucket[16];
    for (size_t e = e0; e < e0 + ne; e++) bucket[b.ent[e].z & 15].push_back(e);
    size_t pos[16] = {}, f = e0;
    while (f < e0 + ne) {
      for (int r = 0; r < 16; r++) {
        if (pos[r] == bucket[r].size()) continue;
        const size_t e = bucket[r][pos[r]++];
        b.fent[f] = b.ent[e];
        for (int t = 0; t < K; t++)
          b.fpx[e0 * K + ((t / 2) * ne + (f - e0)) * 2 + (t & 1)] = b.px[e0 * K + ((t / 2) * ne + (e - e0)) * 2 + (t & 1)];
        for (int c = 0; c < C; c++)
          b.fpy[e0 * C + ((c / 4) * ne + (f - e0)) * 4 + (c & 3)] = b.py[e0 * C + ((c / 4) * ne + (e - e0)) * 4 + (c & 3)];
        f++;
      }
    }
  }
}

}  // namespace icon
