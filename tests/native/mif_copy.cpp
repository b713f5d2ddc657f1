// Reads an image with mrcsd::read_mif and writes it back with mrcsd::write_mif (Float32LE, axis 0
// fastest, or axis 3 fastest with "vc"): the C++ .mif reader and writer the commands use.
#include <cstdio>
#include <cstring>
#include <exception>

#include "host/mif.h"

int main(int argc, char** argv) {
  if (argc < 3) { std::fprintf(stderr, "usage: mif_copy in.mif out.mif [vc]\n"); return 2; }
  try {
    mrcsd::MifImage img = mrcsd::read_mif(argv[1]);
    mrcsd::write_mif(argv[2], img.h, img.data, argc > 3 && std::strcmp(argv[3], "vc") == 0);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mif_copy: %s\n", e.what());
    return 1;
  }
  return 0;
}
