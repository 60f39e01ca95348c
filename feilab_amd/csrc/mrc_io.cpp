// MRC file I/O for the ICON core (host).  See mrc_io.h.
#include "mrc_io.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace mrc {

int element_size(int mode) {
  switch (mode) {
    case kByte: return 1;
    case kShort: return 2;
    case kUShort: return 2;
    case kFloat: return 4;
    default: return 0;
  }
}

bool read_header(FILE* fp, MrcHeader* h) {
  if (std::fseek(fp, 0, SEEK_SET) != 0) return false;
  return std::fread(h, sizeof(MrcHeader), 1, fp) == 1;
}

void init_header(MrcHeader* h, int nx, int ny, int nz) {
  std::memset(h, 0, sizeof(*h));
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->mode = kFloat;
  h->mx = nx; h->my = ny; h->mz = nz;
  h->xlen = (float)nx; h->ylen = (float)ny; h->zlen = (float)nz;
  h->alpha = h->beta = h->gamma = 90.f;
  h->mapc = 1; h->mapr = 2; h->maps = 3;
  h->ispg = 1;
  h->cmap[0] = 'M'; h->cmap[1] = 'A'; h->cmap[2] = 'P'; h->cmap[3] = ' ';
  h->stamp[0] = 'D';
  h->imodStamp = 1146047817;
}

static bool data_offset(FILE* fp, const MrcHeader& h, int64_t* off, std::string* err) {
  const int es = element_size(h.mode);
  if (es == 0) {
    if (err) *err = "unsupported MRC mode " + std::to_string(h.mode);
    return false;
  }
  if (std::fseek(fp, 0, SEEK_END) != 0) return false;
  const int64_t size = (int64_t)ftello(fp);
  const int64_t data = (int64_t)h.nx * h.ny * h.nz * es;
  if (size - data < (int64_t)sizeof(MrcHeader)) {
    if (err) *err = "MRC file shorter than its header says";
    return false;
  }
  *off = size - data;
  return true;
}

bool read_y_slices(FILE* fp, const MrcHeader& h, int y0, int count, float* out, std::string* err) {
  if (y0 < 0 || count < 0 || y0 + count > h.ny) {
    if (err) *err = "slice range outside the tilt series";
    return false;
  }
  int64_t off = 0;
  if (!data_offset(fp, h, &off, err)) return false;
  const int es = element_size(h.mode);
  const int nx = h.nx, nz = h.nz;
  std::vector<unsigned char> buf((size_t)count * nx * es);
  for (int k = 0; k < nz; k++) {
    const int64_t pos = off + ((int64_t)k * h.ny + y0) * (int64_t)nx * es;
    if (fseeko(fp, (off_t)pos, SEEK_SET) != 0 || std::fread(buf.data(), 1, buf.size(), fp) != buf.size()) {
      if (err) *err = "short read in MRC data";
      return false;
    }
    for (int y = 0; y < count; y++) {
      float* dst = out + ((size_t)y * nz + k) * nx;
      const unsigned char* src = buf.data() + (size_t)y * nx * es;
      switch (h.mode) {
        case kByte:
          for (int x = 0; x < nx; x++) dst[x] = (float)src[x];
          break;
        case kShort:
          for (int x = 0; x < nx; x++) { int16_t v; std::memcpy(&v, src + 2 * x, 2); dst[x] = (float)v; }
          break;
        case kUShort:
          for (int x = 0; x < nx; x++) { uint16_t v; std::memcpy(&v, src + 2 * x, 2); dst[x] = (float)v; }
          break;
        default:
          std::memcpy(dst, src, (size_t)nx * 4);
      }
    }
  }
  return true;
}

bool read_z_section(FILE* fp, const MrcHeader& h, int k, float* out, std::string* err) {
  if (k < 0 || k >= h.nz) {
    if (err) *err = "section outside the tilt series";
    return false;
  }
  int64_t off = 0;
  if (!data_offset(fp, h, &off, err)) return false;
  const int es = element_size(h.mode);
  const size_t n = (size_t)h.nx * h.ny;
  std::vector<unsigned char> buf(n * es);
  const int64_t pos = off + (int64_t)k * (int64_t)n * es;
  if (fseeko(fp, (off_t)pos, SEEK_SET) != 0 || std::fread(buf.data(), 1, buf.size(), fp) != buf.size()) {
    if (err) *err = "short read in MRC data";
    return false;
  }
  const unsigned char* src = buf.data();
  switch (h.mode) {
    case kByte:
      for (size_t i = 0; i < n; i++) out[i] = (float)src[i];
      break;
    case kShort:
      for (size_t i = 0; i < n; i++) { int16_t v; std::memcpy(&v, src + 2 * i, 2); out[i] = (float)v; }
      break;
    case kUShort:
      for (size_t i = 0; i < n; i++) { uint16_t v; std::memcpy(&v, src + 2 * i, 2); out[i] = (float)v; }
      break;
    default:
      std::memcpy(out, src, n * 4);
  }
  return true;
}

bool write_real(const char* path, const MrcHeader& in, const float* data, int nx, int ny, std::string* err) {
  MrcHeader h = in;
  h.mz = in.mx;
  h.zlen = in.xlen;
  return write_float(path, h, data, nx, ny, 1, err);
}

bool write_float(const char* path, MrcHeader h, const float* data, int nx, int ny, int nz, std::string* err) {
  h.nx = nx; h.ny = ny; h.nz = nz; h.mode = kFloat;
  if (h.next < 0) h.next = 0;
  // statistics as mrc_update_head computes them
  const size_t plane = (size_t)nx * ny;
  double sum = 0.0, amin = data[0], amax = data[0];
  for (int k = 0; k < nz; k++) {
    double s = 0.0;
    const float* p = data + k * plane;
    for (size_t i = 0; i < plane; i++) {
      if (p[i] > amax) amax = p[i];
      if (p[i] < amin) amin = p[i];
      s += p[i];
    }
    sum += s;
  }
  const double N = (double)nx * ny * nz;
  const double mean = sum / N;
  double var = 0.0;
  for (int k = 0; k < nz; k++) {
    double s = 0.0;
    const float* p = data + k * plane;
    for (size_t i = 0; i < plane; i++) s += (p[i] - mean) * (p[i] - mean);
    var += s;
  }
  h.amin = (float)amin; h.amax = (float)amax; h.amean = (float)mean;
  h.rms = (float)std::sqrt(var / N);
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    if (err) *err = std::string("cannot open ") + path + " for writing";
    return false;
  }
  bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
  if (ok && h.next > 0) {
    std::vector<char> zeros((size_t)h.next, 0);
    ok = std::fwrite(zeros.data(), 1, zeros.size(), fp) == zeros.size();
  }
  if (ok) ok = std::fwrite(data, sizeof(float), (size_t)N, fp) == (size_t)N;
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok && err) *err = std::string("write failed: ") + path;
  return ok;
}

bool write_icon_slice(const char* path, const MrcHeader& in, const float* m, int width, int thickness,
                      std::string* err) {
  MrcHeader h = in;
  h.mz = in.mx;
  h.zlen = in.xlen;
  std::vector<float> t((size_t)width * thickness);
  for (int j = 0; j < thickness; j++)
    for (int i = 0; i < width; i++) t[(size_t)j * width + i] = m[(size_t)i * thickness + j];
  return write_float(path, h, t.data(), width, thickness, 1, err);
}

}  // namespace mrc
