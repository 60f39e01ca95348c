// icon-preprocess: the ICONPreProcess command line
// (sourceCode/ICONPreProcess/ICONPreProcess.c:17-25,90-215): the same flags,
// messages and outputs; the work runs on device 0.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/icon_mi355x.h"

static void help() {
  std::printf("preProcess parameter\n######\n");
  std::printf("-input (-i) : the tilt series.\n######\n");
  std::printf("-tiltfile (-t) : the file containing aligned tilt angle of each projection image. If this option is "
              "not used, then only subtract the mode value of projection images.\n######\n");
  std::printf("-thickness (-th) : the thickness of specimen in pixel. If this option is not used, then only subtract "
              "the mode value of projection images\n######\n");
  std::printf("-output (-o) : preProcessed projection file.\n######\n");
  std::printf("-help (-h) : for help.\n");
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
    std::printf("parameter error!\n input or output is missed.\nPlease use -help to see the manual\n");
    return 0;
  }
  if (!(has_tilt && has_th)) std::printf("tiltfile or thickness is not specified. Normalization will not be executed\n");
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
