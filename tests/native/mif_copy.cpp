// Reads an image with mrcsd::read_mif and writes it back with mrcsd::write_mif (Float32LE, axis 0
// fastest, or axis 3 fastest with "vc"): the C++ .mif reader and writer the commands use.
// "mif_copy --transpose in.mif out.raw": the image's canonical data [nvol][nspatial] turned into
// voxel rows [nspatial][nvol] by mrcsd::transpose_blocked, written as raw float32.
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "host/mif.h"

int main(int argc, char** argv) {
  if (argc < 3) { std::fprintf(stderr, "usage: mif_copy in.mif out.mif [vc]\n"); return 2; }
  try {
    if (std::strcmp(argv[1], "--transpose") == 0) {
      if (argc < 4) { std::fprintf(stderr, "usage: mif_copy --transpose in.mif out.raw\n"); return 2; }
      mrcsd::MifImage img = mrcsd::read_mif(argv[2]);
      int64_t nspatial = 1;
      for (size_t a = 0; a + 1 < img.h.dim.size(); ++a) nspatial *= img.h.dim[a];
      const int64_t nvol = img.h.dim.back();
      std::vector<float> rows(img.data.size());
      mrcsd::transpose_blocked(img.data.data(), nvol, nspatial, rows.data());
      FILE* f = std::fopen(argv[3], "wb");
      if (!f || std::fwrite(rows.data(), sizeof(float), rows.size(), f) != rows.size()) return 1;
      std::fclose(f);
      return 0;
    }
    mrcsd::MifImage img = mrcsd::read_mif(argv[1]);
    mrcsd::write_mif(argv[2], img.h, img.data, argc > 3 && std::strcmp(argv[3], "vc") == 0);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mif_copy: %s\n", e.what());
    return 1;
  }
  return 0;
}
