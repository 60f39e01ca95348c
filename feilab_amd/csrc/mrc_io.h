// MRC file I/O for the ICON core (host).
//
// The on-disk header is the 1024-byte MRC2000/IMOD layout that ICON-GPU reads
// and writes (mrcfile3.h:52-126 in the reference); its field order is fixed by
// the file format.  The data offset is inferred from the file size
// (size - nx*ny*nz*element size), as the reference does (mrcfile3.c:757-783),
// so extended headers of any length are skipped.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

#pragma pack(push, 1)
struct MrcHeader {
  int32_t nx, ny, nz, mode;
  int32_t nxstart, nystart, nzstart;
  int32_t mx, my, mz;
  float xlen, ylen, zlen;
  float alpha, beta, gamma;
  int32_t mapc, mapr, maps;
  float amin, amax, amean;
  int16_t ispg, nsymbt;
  int32_t next;
  int16_t creatid;
  char blank[30];
  int16_t nint, nreal;
  char blank2[20];
  int32_t imodStamp, imodFlags;
  int16_t idtype, lens, nd1, nd2, vd1, vd2;
  float tiltangles[6];
  float xorg, yorg, zorg;
  char cmap[4];
  char stamp[4];
  float rms;
  int32_t nlabl;
  char labels[10][80];
};
#pragma pack(pop)
static_assert(sizeof(MrcHeader) == 1024, "MRC header must be 1024 bytes");

namespace mrc {

enum Mode { kByte = 0, kShort = 1, kFloat = 2, kUShort = 6 };

// Element size of a supported mode, 0 otherwise.
int element_size(int mode);

// Reads the header; false on a short file.
bool read_header(FILE* fp, MrcHeader* h);
// A fresh float-mode header (mrcfile3.c:40-97 defaults).
void init_header(MrcHeader* h, int nx, int ny, int nz);

// Reads rows y0 .. y0+count-1 of every section into out[count][nz][nx]
// (slice-major: out[(y - y0) * nz * nx + k * nx + x]), converting to float.
// This is the reference's 'y'-axis slice read (mrcfile3.c:902-951) for a
// run of slices.  Returns false with *err set on failure.
bool read_y_slices(FILE* fp, const MrcHeader& h, int y0, int count, float* out, std::string* err);

// Section k (nx x ny, the z-axis slice of mrc_read_slice, mrcfile3.c:954-990)
// as float.
bool read_z_section(FILE* fp, const MrcHeader& h, int k, float* out, std::string* err);

// saveMRC_real (saveMRC_real.cu:9-30) followed by mrc_update_head: header
// copied from the tilt series with mz = mx, zlen = xlen, nx, ny, nz = 1, mode
// float; data as given (x fastest).
bool write_real(const char* path, const MrcHeader& in, const float* data, int nx, int ny, std::string* err);

// Writes a float image nx (fast) x ny x nz with header `h` (nx, ny, nz, mode
// overwritten), `h.next` zero bytes of extended header, then the data, and
// fills amin/amax/amean/rms as mrc_update_head (mrcfile3.c:113-514) computes
// them: double sums per section, mean = sum/N, rms = sqrt(sum (v-mean)^2 / N).
bool write_float(const char* path, MrcHeader h, const float* data, int nx, int ny, int nz, std::string* err);

// The per-slice output of ICON (saveMRC_real_inv, saveMRC_real.cu:33-56):
// header copied from the tilt series with mz = mx, zlen = xlen, nx = width,
// ny = thickness, nz = 1, mode = float; image transposed so x is fastest.
// m is the solver's [width][thickness] array.
bool write_icon_slice(const char* path, const MrcHeader& in, const float* m, int width, int thickness,
                      std::string* err);

}  // namespace mrc
