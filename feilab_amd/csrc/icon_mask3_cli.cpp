// icon-mask3: the ICONMask3 command line
// (ICONCrossVMaskIn2D_weightOnlyOnSampled.cu:26-42,602-800): the same flags
// and defaults; the filtering runs on device 0.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/icon_mi355x.h"

static void help() {
  std::printf(
      "usage: icon-mask3 -i DIR -s FIRST,LAST -th ROWS -t TILTS -o OUT.mrc\n"
      "                  [-r PX | -cf CROSSV.frc -ff FULLREC.frc] [-gw PX] [-z PX] [-nm 0|1]\n"
      "Stacks the slices DIR/midYYYYY.mrc (ICON's reconstruction folder) into one volume,\n"
      "each slice filtered in 2D Fourier space by a mask built from the tilt range and\n"
      "the resolution the cross-validation FRC supports.\n"
      "  -inputPath, -i   folder holding the reconstructed slices midYYYYY.mrc\n"
      "  -slice, -s       first and last slice to stack, inclusive (\"0,511\": 512 slices)\n"
      "  -thickness, -th  rows kept around the slice centre (the output's depth, pixels)\n"
      "  -tiltfile, -t    tilt angles of the series (degrees, as aligned)\n"
      "  -output, -o      the stacked, filtered volume (MRC)\n"
      "  -radius, -r      Fourier mask radius in pixels; replaces the two FRC files\n"
      "  -crossVfrc, -cf  FRC of the cross-validation reprojection (ICON's crossV.frc)\n"
      "  -fullRecfrc, -ff FRC of the full reconstruction's reprojection (fullRec.frc)\n"
      "  -gaussWidth, -gw width in pixels of the mask's Gaussian edge (default: the gap\n"
      "                   between the FRC-0.5 and FRC-0.3 radii, or 10 with -radius)\n"
      "  -zshift, -z      depth offset of the kept rows in pixels (default 0)\n"
      "  -noMask, -nm     1: stack the cropped slices without filtering (default 0)\n"
      "  -help, -h        this text\n");
}

int main(int argc, char** argv) {
  std::string in, tilt, out, cf, ff;
  int s0 = 0, s1 = -1, thickness = 0, radius = 0, zshift = 0, gw = -1, no_mask = 0;
  bool m[11] = {false};
  for (int i = 1; i < argc;) {
    if (argv[i][0] != '-') { i++; continue; }
    const char* k = argv[i] + 1;
    const char* v = i + 1 < argc ? argv[i + 1] : "";
    if (!std::strcmp(k, "inputPath") || !std::strcmp(k, "i")) { in = v; m[0] = true; i += 2; }
    else if (!std::strcmp(k, "slice") || !std::strcmp(k, "s")) { std::sscanf(v, "%d,%d", &s0, &s1); m[1] = true; i += 2; }
    else if (!std::strcmp(k, "thickness") || !std::strcmp(k, "th")) { std::sscanf(v, "%d", &thickness); m[2] = true; i += 2; }
    else if (!std::strcmp(k, "radius") || !std::strcmp(k, "r")) { std::sscanf(v, "%d", &radius); m[3] = true; i += 2; }
    else if (!std::strcmp(k, "tiltfile") || !std::strcmp(k, "t")) { tilt = v; m[4] = true; i += 2; }
    else if (!std::strcmp(k, "output") || !std::strcmp(k, "o")) { out = v; m[5] = true; i += 2; }
    else if (!std::strcmp(k, "crossVfrc") || !std::strcmp(k, "cf")) { cf = v; m[6] = true; i += 2; }
    else if (!std::strcmp(k, "fullRecfrc") || !std::strcmp(k, "ff")) { ff = v; m[7] = true; i += 2; }
    else if (!std::strcmp(k, "zshift") || !std::strcmp(k, "z")) { std::sscanf(v, "%d", &zshift); m[8] = true; i += 2; }
    else if (!std::strcmp(k, "gaussWidth") || !std::strcmp(k, "gw")) { std::sscanf(v, "%d", &gw); m[9] = true; i += 2; }
    else if (!std::strcmp(k, "noMask") || !std::strcmp(k, "nm")) { std::sscanf(v, "%d", &no_mask); m[10] = true; i += 2; }
    else if (!std::strcmp(k, "help") || !std::strcmp(k, "h")) { help(); return 1; }
    else i++;
  }
  std::printf("parameter:\ninputPath : %s\nslice : %d,%d\nthickness : %d\n", in.c_str(), s0, s1, thickness);
  if (m[3]) std::printf("radius : %d\n", radius);
  std::printf("zshift : %d\ntiltfile : %s\noutput : %s\n", zshift, tilt.c_str(), out.c_str());
  if (!m[3] && m[6] && m[7]) std::printf("crossVfrc : %s\nfullRecfrc : %s\n", cf.c_str(), ff.c_str());
  if (m[9]) std::printf("gaussWidth : %d\n", gw);
  std::printf("noMask : %d\n", no_mask);
  if (!(m[0] && m[1] && m[2] && m[4] && m[5])) {
    std::printf("parameter error!\nneeded: -inputPath, -slice, -thickness, -tiltfile and -output "
                "(-help lists the options)\n");
    return -1;
  }
  if (no_mask == 1) std::printf("-noMask 1: slices are cropped and stacked without the Fourier mask\n");
  if (no_mask == 0 && !m[3] && !(m[6] && m[7])) {
    std::printf("parameter error!\nthe mask needs -radius, or both -crossVfrc and -fullRecfrc "
                "(-help lists the options)\n");
    return -1;
  }
  const int rc = icon_mask3(0, in.c_str(), s0, s1, thickness, tilt.c_str(), out.c_str(), m[3] ? radius : 0,
                            m[6] ? cf.c_str() : nullptr, m[7] ? ff.c_str() : nullptr, zshift, m[9] ? gw : -1,
                            no_mask == 1 ? 1 : 0);
  if (rc != ICON_OK) {
    std::printf("%s\n", icon_mask3_error());
    return 1;
  }
  return no_mask == 1 ? 1 : 0;
}
