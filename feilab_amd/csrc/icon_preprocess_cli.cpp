// icon-preprocess: the ICONPreProcess command line
// (sourceCode/ICONPreProcess/ICONPreProcess.c:17-25,90-215): the same flags,
// defaults, exit codes and outputs; the work runs on device 0.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/icon_mi355x.h"

static void help() {
  std::printf(
      "usage: icon-preprocess -i IN.mrc -o OUT.mrc [-t TILTS -th THICKNESS]\n"
      "Subtracts from every projection the mode of its central region's histogram and,\n"
      "given the tilt angles and the specimen thickness, scales each projection to the\n"
      "variance 0.33 * thickness / cos(tilt).\n"
      "  -input, -i      tilt series to prepare (MRC)\n"
      "  -output, -o     prepared tilt series (MRC, float)\n"
      "  -tiltfile, -t   aligned tilt angle of every projection (degrees); without it\n"
      "                  only the mode is subtracted\n"
      "  -thickness, -th specimen thickness in pixels; without it only the mode is\n"
      "                  subtracted\n"
      "  -help, -h       this text\n");
}

int main(int argc, char** argv) {
  std::string in, tilt, out;
  int thickness = 0;
  bool has_in = false, has_tilt = false, has_out = false, has_th = false;
  for (int i = 1; i < argc;) {
    if (argv[i][0] != '-') { i++; continue; }
    const char* k = argv[i] + 1;
    auto val = [&]() -> const char* { return i + 1 < argc ? argv[i + 1] : ""; };
    if (!std::strcmp(k, "input") || !std::strcmp(k, "i")) { in = val(); has_in = true; i += 2; }
    else if (!std::strcmp(k, "tiltfile") || !std::strcmp(k, "t")) { tilt = val(); has_tilt = true; i += 2; }
    else if (!std::strcmp(k, "output") || !std::strcmp(k, "o")) { out = val(); has_out = true; i += 2; }
    else if (!std::strcmp(k, "thickness") || !std::strcmp(k, "th")) {
      std::sscanf(val(), "%d", &thickness);
      has_th = true;
      i += 2;
    } else if (!std::strcmp(k, "help") || !std::strcmp(k, "h")) {
      help();
      return 0;
    } else {
      i++;
    }
  }
  if (!(has_in && has_out)) {
    std::printf("parameter error!\nneeded: -input and -output (-help lists the options)\n");
    return 0;
  }
  if (!(has_tilt && has_th)) std::printf("no tilt file or thickness: the mode is subtracted, the variance is left as is\n");
  std::printf("parameter:\n");
  std::printf("input : %s\n", in.c_str());
  if (has_tilt) std::printf("tiltfile : %s\n", tilt.c_str());
  if (has_th) std::printf("thickness : %d\n", thickness);
  std::printf("output : %s\n", out.c_str());
  const bool norm = has_tilt && has_th;
  const int rc = icon_preprocess_file(0, in.c_str(), out.c_str(), norm ? tilt.c_str() : nullptr, norm ? thickness : 0);
  if (rc != ICON_OK) {
    std::printf("\n%s\n", icon_preprocess_error());
    return 1;
  }
  return 0;
}
