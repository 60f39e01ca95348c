// icon-gpu: command line of the MI355X ICON core, with ICON-GPU's flags,
// defaults and exit codes (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/
// ICONGPU.cu:93-609):
//   icon-gpu -i <tilts.mrc> -t <angles.tlt> -o <dir> -s y0,y1 -iter a,b,c
//            [-d 1|2] [-thr f] [-g i,j,...|-1] [-thi T] [-skipCV 0|1] [-h]
// Output: <dir>/reconstruction/midYYYYY.mrc per slice (and <dir>/crossValidation).
// Slices are split over the listed devices in proportion to their compute
// units (ICONGPU.cu:393-431), one host thread per device.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/icon_mi355x.h"
#include "mrc_io.h"


namespace {

void help() {
  std::printf("ICON parameter\n######\n");
  std::printf("-input (-i) : the aligned tilt series.\n######\n");
  std::printf("-tiltfile (-t) : the aligned tilt file.\n######\n");
  std::printf("-outputPath (-o) : the result folder; 'crossValidation' and 'reconstruction' are created inside.\n######\n");
  std::printf("-slice (-s) : first,last slice to reconstruct, e.g. 0,511.\n######\n");
  std::printf("-ICONIteration (-iter) : INFR,ICON,INFR iteration counts, e.g. 5,50,10.\n######\n");
  std::printf("-dataType (-d) : 1 cryoET / plastic (signal black), 2 negative stain (signal white); default 1.\n######\n");
  std::printf("-threshold (-thr) : ICON threshold, default 0.03 (dataType 1) / -0.03 (dataType 2).\n######\n");
  std::printf("-thickness (-thi) : reconstruction thickness in pixels, default the width.\n######\n");
  std::printf("-skipCrossValidation (-skipCV) : 1 skips the cross validation, default 0.\n######\n");
  std::printf("-gpu (-g) : device list, e.g. 0,2; default -1 = all devices.\n######\n");
  std::printf("-help (-h) : for help.\n");
}

}  // namespace

// The cross-validation outputs (ICONGPU.cu:468-592): the measured projection
// at the least-|angle| tilt for the reconstructed slices (GroundTruth.mrc),
// the reprojections of the full and the cross-validation solves, and their
// smoothed FRCs against the measurement (fullRec.frc, crossV.frc: "%f %f"
// lines of spatial frequency i/nx and FRC, i < nx/2).
static bool cross_validation_outputs(const std::string& input, const std::string& tilt, const std::string& cvdir,
                                     const MrcHeader& head, int y0, int total, const std::vector<float>& full,
                                     const std::vector<float>& cv) {
  const int nx = head.nx;
  std::vector<float> tl(1000);
  int nt = 0;
  if (icon_read_tilts(tilt.c_str(), tl.data(), (int)tl.size(), &nt) != ICON_OK || nt < 1) {
    std::printf("cannot read %s\n", tilt.c_str());
    return false;
  }
  float om = -tl[0];
  int omit = 0;
  for (int i = 0; i < nt; i++)
    if (std::fabs(om) > std::fabs(-tl[i])) { om = -tl[i]; omit = i; }
  std::vector<float> sec((size_t)nx * head.ny), gt((size_t)total * nx);
  std::string err;
  FILE* fp = std::fopen(input.c_str(), "rb");
  const bool ok = fp && mrc::read_z_section(fp, head, omit, sec.data(), &err);
  if (fp) std::fclose(fp);
  if (!ok) {
    std::printf("cannot read the projection at tilt %d: %s\n", omit, err.c_str());
    return false;
  }
  for (int r = 0; r < total; r++)
    std::memcpy(&gt[(size_t)r * nx], &sec[(size_t)(y0 + r) * nx], sizeof(float) * nx);
  const struct { const char* name; const float* data; } files[] = {
      {"GroundTruth.mrc", gt.data()}, {"fullRec_reProjection.mrc", full.data()}, {"crossV_reprojection.mrc", cv.data()}};
  for (const auto& f : files) {
    const std::string path = cvdir + "/" + f.name;
    if (!mrc::write_real(path.c_str(), head, f.data, nx, total, &err)) {
      std::printf("cannot write %s: %s\n", path.c_str(), err.c_str());
      return false;
    }
  }
  const double resstep = 1.0 / nx;
  const struct { const char* name; const float* rec; } curves[] = {{"fullRec.frc", full.data()}, {"crossV.frc", cv.data()}};
  std::vector<float> frc(nx / 2 + 1);
  for (const auto& c : curves) {
    if (icon_frc(gt.data(), c.rec, nx, total, frc.data()) != ICON_OK) return false;
    icon_frc_smooth(frc.data(), nx / 2, 10);
    const std::string path = cvdir + "/" + c.name;
    FILE* out = std::fopen(path.c_str(), "w");
    if (!out) {
      std::printf("cannot write %s\n", path.c_str());
      return false;
    }
    for (int i = 0; i < nx / 2; i++) std::fprintf(out, "%f %f\n", resstep * i, frc[i]);
    std::fclose(out);
  }
  return true;
}

int main(int argc, char** argv) {
  std::string input, tilt, out;
  int y0 = 0, y1 = -1, pre = 0, mid = 0, post = 0, data_type = 1, thickness = 0, skip_cv = 0;
  float thr = 520.f;
  bool have[5] = {false, false, false, false, false};
  std::vector<int> devices;
  int ndev = 0;
  if (icon_device_count(&ndev) != ICON_OK) ndev = 0;
  std::printf("##################\n%d GPU(s) are detected in the system!\n", ndev);
  for (int d = 0; d < ndev; d++) devices.push_back(d);
  for (int i = 1; i < argc;) {
    const char* a = argv[i];
    if (a[0] != '-') { i++; continue; }
    const char* k = a + 1;
    auto val = [&]() -> const char* { return (i + 1 < argc) ? argv[i + 1] : ""; };
    if (!std::strcmp(k, "input") || !std::strcmp(k, "i")) { input = val(); have[0] = true; i += 2; }
    else if (!std::strcmp(k, "tiltfile") || !std::strcmp(k, "t")) { tilt = val(); have[1] = true; i += 2; }
    else if (!std::strcmp(k, "outputPath") || !std::strcmp(k, "o")) { out = val(); have[2] = true; i += 2; }
    else if (!std::strcmp(k, "slice") || !std::strcmp(k, "s")) {
      if (std::sscanf(val(), "%d,%d", &y0, &y1) == 2) have[3] = true;
      i += 2;
    } else if (!std::strcmp(k, "ICONIteration") || !std::strcmp(k, "iter")) {
      if (std::sscanf(val(), "%d,%d,%d", &pre, &mid, &post) == 3) have[4] = true;
      i += 2;
    } else if (!std::strcmp(k, "dataType") || !std::strcmp(k, "d")) { std::sscanf(val(), "%d", &data_type); i += 2; }
    else if (!std::strcmp(k, "threshold") || !std::strcmp(k, "thr")) { std::sscanf(val(), "%f", &thr); i += 2; }
    else if (!std::strcmp(k, "thickness") || !std::strcmp(k, "thi")) { std::sscanf(val(), "%d", &thickness); i += 2; }
    else if (!std::strcmp(k, "skipCrossValidation") || !std::strcmp(k, "skipCV")) {
      std::sscanf(val(), "%d", &skip_cv);
      i += 2;
    } else if (!std::strcmp(k, "gpu") || !std::strcmp(k, "g")) {
      std::vector<int> list;
      std::string s = val();
      size_t p = 0;
      bool all = false;
      while (p <= s.size()) {
        size_t q = s.find(',', p);
        if (q == std::string::npos) q = s.size();
        int d = std::atoi(s.substr(p, q - p).c_str());
        if (d == -1) { all = true; break; }
        list.push_back(d);
        p = q + 1;
      }
      if (!all && !list.empty()) devices = list;
      i += 2;
    } else if (!std::strcmp(k, "help") || !std::strcmp(k, "h")) {
      help();
      return -1;
    } else {
      i++;
    }
  }
  for (bool h : have)
    if (!h) {
      std::printf("parameter error!\n Please use -help to see the manual\n");
      return -1;
    }
  // threshold sentinel (ICONGPU.cu:317-323)
  if (std::fabs(thr - 520.f) < 0.0001f) thr = data_type == 1 ? 0.03f : (data_type == 2 ? -0.03f : thr);
  const int iternum = pre + mid + post;
  const int total = y1 - y0 + 1;
  std::printf("parameter:\ninput : %s\ntiltfile : %s\noutputPath : %s\nslice : %d,%d\n", input.c_str(),
              tilt.c_str(), out.c_str(), y0, y1);
  std::printf("ICONIteration : %d,%d,%d\ndataType : %d\nthreshold : %f\n", pre, mid, post, data_type, thr);
  if (thickness != 0) std::printf("thickness : %d\n", thickness);
  const int used_n = (int)devices.size() < total ? (int)devices.size() : total;
  if (used_n <= 0) {
    std::printf("Notice : No GPU is specified to use!\n");
    return 0;
  }
  if (skip_cv == 1) std::printf("skip cross validation!\n");
  const std::string cvdir = out + "/crossValidation", recdir = out + "/reconstruction";
  mkdir(cvdir.c_str(), 0755);
  mkdir(recdir.c_str(), 0755);
  std::vector<int> cus(devices.size(), 1), first(devices.size()), last(devices.size());
  for (size_t d = 0; d < devices.size(); d++) icon_device_compute_units(devices[d], &cus[d]);
  const int used = icon_partition_slices((int)devices.size(), cus.data(), y0, y1, first.data(), last.data());
  int before[5] = {pre, post, 1, data_type, skip_cv};
  float betain[2] = {thr, (float)thickness};
  MrcHeader head;
  {
    FILE* fp = std::fopen(input.c_str(), "rb");
    if (!fp || !mrc::read_header(fp, &head)) {
      std::printf("cannot read %s\n", input.c_str());
      if (fp) std::fclose(fp);
      return 1;
    }
    std::fclose(fp);
  }
  std::vector<float> reproj((size_t)total * head.nx, 0.f), reproj_cv((size_t)total * head.nx, 0.f);
  std::vector<int> status(used, 0);
  std::vector<std::thread> th;
  for (int d = 0; d < used; d++) {
    std::printf("thread %d using GPU %d , deals with slices : start %d end %d\n", d, devices[d], first[d], last[d]);
    th.emplace_back([&, d]() {
      const size_t off = (size_t)(first[d] - y0) * head.nx;
      status[d] = icon_lib_2d2(devices[d], first[d], last[d] + 1, iternum, input.c_str(), reproj.data() + off,
                               tilt.c_str(), recdir.c_str(), before, betain, 0);
      if (status[d] == ICON_OK && skip_cv == 0)   // ICONGPU.cu:52-54
        status[d] = icon_lib_2d_crossv(devices[d], first[d], last[d] + 1, iternum, input.c_str(),
                                       reproj_cv.data() + off, tilt.c_str(), recdir.c_str(), before, betain, 0);
    });
  }
  for (auto& t : th) t.join();
  for (int d = 0; d < used; d++)
    if (status[d] != ICON_OK) {
      std::printf("device %d failed: %s\n", devices[d], icon_lib_2d2_error());
      return 1;
    }
  if (skip_cv == 0 && !cross_validation_outputs(input, tilt, cvdir, head, y0, total, reproj, reproj_cv)) return 1;
  return 0;
}
