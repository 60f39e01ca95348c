// File-level drop-in of ICON-GPU's per-device worker Lib_2D2
// (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/MainCode.cu:3-177):
// angles -> plan -> for each slice: read the sinogram row of every tilt,
// solve, write <result_dir>/midYYYYY.mrc and the reprojection row.
// Differences: slices are solved in batches on the GPU, errors are returned
// as status codes instead of exit(1), and the tilt count must match the
// stack's nz.
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icon_mi355x.h"
#include "mrc_io.h"


namespace {
std::mutex g_lmu;
std::string g_lerr_store;
void set_err(const std::string& s) {
  std::lock_guard<std::mutex> lk(g_lmu);
  g_lerr_store = s;
  std::fprintf(stderr, "icon_lib_2d2: %s\n", s.c_str());
}
}  // namespace


extern "C" int icon_read_tilts(const char* path, float* tilt_deg, int cap, int* count) {
  // MainCode.cu:38-45: whitespace-separated degrees (negation happens in the plan)
  FILE* fp = std::fopen(path, "r");
  if (!fp) return ICON_ERR_IO;
  int n = 0;
  float v;
  while (std::fscanf(fp, "%f", &v) == 1) {
    if (n < cap) tilt_deg[n] = v;
    n++;
  }
  std::fclose(fp);
  *count = n;
  return n > cap ? ICON_ERR_CAPACITY : ICON_OK;
}

extern "C" int icon_mrc_read_y_slices(const char* path, int y0, int count, float* out, int* dims4) {
  if (!path) return ICON_ERR_ARG;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) {
    set_err(std::string("cannot open ") + path);
    return ICON_ERR_IO;
  }
  MrcHeader h;
  std::string err;
  int rc = ICON_OK;
  if (!mrc::read_header(fp, &h)) {
    set_err("cannot read MRC header");
    rc = ICON_ERR_IO;
  } else {
    if (dims4) { dims4[0] = h.nx; dims4[1] = h.ny; dims4[2] = h.nz; dims4[3] = h.mode; }
    if (out && count > 0 && !mrc::read_y_slices(fp, h, y0, count, out, &err)) {
      set_err(err);
      rc = ICON_ERR_IO;
    }
  }
  std::fclose(fp);
  return rc;
}

extern "C" int icon_mrc_write_slice(const char* path, const char* template_path, const float* m, int width,
                                    int thickness) {
  if (!path || !template_path || !m) return ICON_ERR_ARG;
  FILE* fp = std::fopen(template_path, "rb");
  if (!fp) {
    set_err(std::string("cannot open ") + template_path);
    return ICON_ERR_IO;
  }
  MrcHeader h;
  const bool ok = mrc::read_header(fp, &h);
  std::fclose(fp);
  if (!ok) {
    set_err("cannot read MRC header");
    return ICON_ERR_IO;
  }
  std::string err;
  if (!mrc::write_icon_slice(path, h, m, width, thickness, &err)) {
    set_err(err);
    return ICON_ERR_IO;
  }
  return ICON_OK;
}

extern "C" const char* icon_lib_2d2_error(void) {
  std::lock_guard<std::mutex> lk(g_lmu);
  return g_lerr_store.c_str();
}

// Slice partition across devices (ICONGPU.cu:399-453): shares proportional to
// the compute-unit count (floored), leftovers handed out round-robin in the
// order of most units first; the last device takes whatever remains.  Ties in
// that order resolve as the reference's selection loop does (ICONGPU.cu:
// 402-414): each pick starts from the last device not yet ordered and moves
// only to a device of strictly more units, scanning from device 0, so equal
// devices are ordered last index first.
extern "C" int icon_partition_slices(int ndev, const int* cus, int y_start, int y_end, int* first, int* last) {
  const int total = y_end - y_start + 1;
  if (ndev < 1 || total < 1) return ICON_ERR_ARG;
  std::vector<int> rate(ndev, 0), order(ndev, 0), taken(ndev, 0);
  for (int d = 0; d < ndev; d++) {
    int pick = -1;
    for (int i = ndev - 1; i >= 0 && pick < 0; i--)
      if (!taken[i]) pick = i;
    for (int j = 0; j < ndev; j++)
      if (!taken[j] && cus[pick] < cus[j]) pick = j;
    order[d] = pick;
    taken[pick] = 1;
  }
  double sum = 0;
  for (int d = 0; d < ndev; d++) sum += cus[d];
  for (int d = 0; d < ndev; d++) rate[d] = (int)std::floor((double)cus[d] / sum * total);
  int assigned = 0;
  for (int d = 0; d < ndev; d++) assigned += rate[d];
  for (int i = 0, j = 0; i < total - assigned; i++) {
    rate[order[j]]++;
    j = (j + 1) % ndev;
  }
  const int used = ndev < total ? ndev : total;
  int next = 0;
  for (int d = 0; d < used; d++) {
    int s, e;
    if (d != ndev - 1) {
      s = next;
      e = rate[d] != 0 ? s + rate[d] - 1 : s;
      next = e + 1;
    } else {
      s = next;
      e = y_end - y_start;
    }
    first[d] = s + y_start;
    last[d] = e + y_start;
  }
  return used;
}



// Lib_2D2 (cv = false) and Lib_2D_crossV2 (cv = true): the same batched
// solve, the cross-validation one on a plan without the least-|angle| tilt,
// writing no slice files.
static int lib_2d(bool cv, int device, int slice_begin, int slice_end, int iternum, const char* mrc_path,
                  float* reproj, const char* angle_path, const char* result_dir, const int* before,
                  const float* betain, int batch) {
  if (!mrc_path || !angle_path || !result_dir || !before || !betain) {
    set_err(cv ? "icon_lib_2d_crossv: null argument" : "icon_lib_2d2: null argument");
    return ICON_ERR_ARG;
  }
  FILE* fin = std::fopen(mrc_path, "rb");
  if (!fin) {
    set_err(std::string("cannot open tilt series ") + mrc_path);
    return ICON_ERR_IO;
  }
  MrcHeader head;
  std::string err;
  if (!mrc::read_header(fin, &head)) {
    std::fclose(fin);
    set_err("cannot read MRC header");
    return ICON_ERR_IO;
  }
  const int nx = head.nx, nz = head.nz;
  std::vector<float> tilts(1000);
  int ntilt = 0;
  int rc = icon_read_tilts(angle_path, tilts.data(), (int)tilts.size(), &ntilt);
  if (rc == ICON_ERR_IO) {
    std::fclose(fin);
    set_err(std::string("Can not open anglefile '") + angle_path + "'!");
    return ICON_ERR_IO;
  }
  if (ntilt < nz) {
    std::fclose(fin);
    set_err("tilt file has fewer angles than the stack has sections");
    return ICON_ERR_ARG;
  }
  const int pre = before[0], post = before[1], data_type = before[3];
  const float thr = betain[0];
  const int thickness = (int)betain[1] > 0 ? (int)betain[1] : nx;
  const int icon = iternum - pre - post;
  if (icon < 0) {
    std::fclose(fin);
    set_err("iteration split exceeds the iteration count");
    return ICON_ERR_ARG;
  }
  if (slice_end > head.ny) slice_end = head.ny;
  const int total = slice_end - slice_begin;
  if (total <= 0) {
    std::fclose(fin);
    return ICON_OK;
  }
  const int B = batch > 0 ? (batch < total ? batch : total) : (total < 256 ? total : 256);
  icon_plan* plan = nullptr;
  rc = cv ? icon_plan_create_cv(&plan, device, nx, thickness, nz, tilts.data(), B)
          : icon_plan_create(&plan, device, nx, thickness, nz, tilts.data(), B);
  if (rc) {
    std::fclose(fin);
    set_err(icon_last_error());
    return rc;
  }
  // The solve of each batch is replayed as a HIP graph on device buffers
  // allocated once for the whole slice range (the last, shorter batch
  // captures its own graph).
  icon_solve_opts o{};
  o.pre = pre; o.icon = icon; o.post = post; o.data_type = data_type; o.threshold = thr; o.use_graph = 1;
  std::vector<float> sino((size_t)B * nz * nx), m((size_t)B * nx * thickness), rp((size_t)B * nx);
  void *d_sino = nullptr, *d_m = nullptr, *d_rp = nullptr;
  int prev_dev = -1;
  if (icon_get_device(&prev_dev) != ICON_OK) prev_dev = -1;
  if (icon_set_device(device) != ICON_OK || icon_malloc(&d_sino, sizeof(float) * sino.size()) != ICON_OK ||
      icon_malloc(&d_m, sizeof(float) * m.size()) != ICON_OK || icon_malloc(&d_rp, sizeof(float) * rp.size()) != ICON_OK) {
    rc = ICON_ERR_DEVICE;
    set_err(icon_last_error());
  }
  for (int y0 = slice_begin; rc == ICON_OK && y0 < slice_end; y0 += B) {
    const int cnt = (slice_end - y0) < B ? (slice_end - y0) : B;
    if (!mrc::read_y_slices(fin, head, y0, cnt, sino.data(), &err)) {
      rc = ICON_ERR_IO;
      set_err(err);
      break;
    }
    auto t0 = std::chrono::steady_clock::now();
    rc = icon_memcpy_h2d(d_sino, sino.data(), sizeof(float) * (size_t)cnt * nz * nx);
    if (!rc) rc = icon_reconstruct(plan, cnt, d_sino, d_m, nullptr, d_rp, nullptr, nullptr, &o, nullptr);
    if (!rc) rc = icon_device_synchronize();
    if (!rc) rc = icon_memcpy_d2h(m.data(), d_m, sizeof(float) * (size_t)cnt * nx * thickness);
    if (!rc) rc = icon_memcpy_d2h(rp.data(), d_rp, sizeof(float) * (size_t)cnt * nx);
    if (rc) {
      set_err(icon_last_error());
      break;
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int s = 0; s < cnt; s++) {
      const int Y = y0 + s;
      if (cv) {
        std::printf("cross validation!\n");
      } else {
        std::printf("slice %d finished, time %lf\n", Y, dt / cnt);
        char path[4096];
        std::snprintf(path, sizeof path, "%s/mid%05d.mrc", result_dir, Y);
        if (!mrc::write_icon_slice(path, head, m.data() + (size_t)s * nx * thickness, nx, thickness, &err)) {
          rc = ICON_ERR_IO;
          set_err(err);
          break;
        }
      }
      if (reproj) std::memcpy(reproj + (size_t)(Y - slice_begin) * nx, rp.data() + (size_t)s * nx, sizeof(float) * nx);
    }
    if (rc) break;
  }
  if (!cv) std::printf("full reconstruction\n");
  icon_free(d_sino);
  icon_free(d_m);
  icon_free(d_rp);
  icon_plan_destroy(plan);
  if (prev_dev >= 0) icon_set_device(prev_dev);
  std::fclose(fin);
  return rc;
}

extern "C" int icon_lib_2d2(int device, int slice_begin, int slice_end, int iternum, const char* mrc_path,
                            float* reproj, const char* angle_path, const char* result_dir, const int* before,
                            const float* betain, int batch) {
  return lib_2d(false, device, slice_begin, slice_end, iternum, mrc_path, reproj, angle_path, result_dir, before,
                betain, batch);
}

extern "C" int icon_lib_2d_crossv(int device, int slice_begin, int slice_end, int iternum, const char* mrc_path,
                                  float* reproj, const char* angle_path, const char* result_dir, const int* before,
                                  const float* betain, int batch) {
  return lib_2d(true, device, slice_begin, slice_end, iternum, mrc_path, reproj, angle_path, result_dir, before,
                betain, batch);
}

// ---------------------------------------------------------------- FRC
// calFRC_nonsqure (calFRC.cu:93-174): ring r = floor(sqrt(di^2 + dj^2 (nx/ny)^2))
// of the centred nx x ny spectrum, r < nx/2; frc[0] = 1.  The 2D transforms
// are separable DFTs in double on the host (any nx, ny).
static void dft2(const float* img, int nx, int ny, std::vector<double>& re, std::vector<double>& im) {
  std::vector<double> cx(nx), sx(nx), cy(ny), sy(ny);
  for (int k = 0; k < nx; k++) { cx[k] = std::cos(2 * M_PI * k / nx); sx[k] = -std::sin(2 * M_PI * k / nx); }
  for (int k = 0; k < ny; k++) { cy[k] = std::cos(2 * M_PI * k / ny); sy[k] = -std::sin(2 * M_PI * k / ny); }
  std::vector<double> tr((size_t)nx * ny), ti((size_t)nx * ny);
  for (int j = 0; j < ny; j++)          // rows: transform along x
    for (int u = 0; u < nx; u++) {
      double ar = 0, ai = 0;
      for (int i = 0; i < nx; i++) {
        const int q = (int)(((long)u * i) % nx);
        const double v = img[(size_t)j * nx + i];
        ar += v * cx[q];
        ai += v * sx[q];
      }
      tr[(size_t)j * nx + u] = ar;
      ti[(size_t)j * nx + u] = ai;
    }
  re.assign((size_t)nx * ny, 0.0);
  im.assign((size_t)nx * ny, 0.0);
  for (int v = 0; v < ny; v++)          // columns: transform along y
    for (int j = 0; j < ny; j++) {
      const int q = (int)(((long)v * j) % ny);
      const double c = cy[q], sn = sy[q];
      const double* a = &tr[(size_t)j * nx];
      const double* b = &ti[(size_t)j * nx];
      double* orr = &re[(size_t)v * nx];
      double* oi = &im[(size_t)v * nx];
      for (int u = 0; u < nx; u++) {
        orr[u] += a[u] * c - b[u] * sn;
        oi[u] += a[u] * sn + b[u] * c;
      }
    }
}

extern "C" int icon_frc(const float* image1, const float* image2, int nx, int ny, float* frc) {
  if (!image1 || !image2 || !frc || nx < 2 || ny < 1) {
    set_err("icon_frc: bad argument");
    return ICON_ERR_ARG;
  }
  std::vector<double> r1, i1, r2, i2;
  dft2(image1, nx, ny, r1, i1);
  dft2(image2, nx, ny, r2, i2);
  const int h = nx / 2, hy = ny / 2;
  std::vector<double> up(h + 1, 0.0), l1(h + 1, 0.0), l2(h + 1, 0.0);
  const float damp2 = ((float)nx / (float)ny) * ((float)nx / (float)ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int i1c = (i + h) % nx, j1c = (j + hy) % ny;
      // float arithmetic as the reference: int^2 + (float)(int^2) * damp2
      const float q = (float)((i1c - h) * (i1c - h)) + (float)((j1c - hy) * (j1c - hy)) * damp2;
      const int r = (int)std::floor(std::sqrt((double)q));
      if (r < h) {
        const size_t k = (size_t)j * nx + i;
        up[r] += r1[k] * r2[k] + i1[k] * i2[k];
        l1[r] += r1[k] * r1[k] + i1[k] * i1[k];
        l2[r] += r2[k] * r2[k] + i2[k] * i2[k];
      }
    }
  frc[0] = 1.f;
  for (int r = 1; r < h; r++) {
    const double d = std::sqrt(l1[r]) * std::sqrt(l2[r]);
    frc[r] = d != 0 ? (float)(up[r] / d) : 0.f;
  }
  return ICON_OK;
}

// frcSmooth (ICONGPU.cu:59-91): running mean over a window of `overlap`
// entries, shifted inwards at the ends.
extern "C" void icon_frc_smooth(float* frc, int size, int overlap) {
  std::vector<float> t(size);
  for (int i = 0; i < size; i++) {
    int jb = i - overlap / 2;
    if (jb < 0) jb = 0;
    int je = jb + overlap;
    if (je > size - 1) {
      jb -= je - size + 1;
      je = size - 1;
    }
    if (jb < 0) jb = 0;
    float sum = 0.f;
    int num = 0;
    for (int j = jb; j < je; j++) {
      sum += frc[j];
      num++;
    }
    t[i] = num > 0 ? sum / num : frc[i];
  }
  for (int i = 0; i < size; i++) frc[i] = t[i];
}
