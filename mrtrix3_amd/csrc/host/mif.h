// Minimal MRtrix .mif / .mih image I/O for the dwi2fod drop-in (core/formats/mrtrix.cpp:40-117,
// core/formats/mrtrix_utils.h:64-182): text header "mrtrix image", key: value lines, "END",
// data at the byte offset given by "file: . <offset>" (or a separate file for .mih).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace mrcsd {

struct MifHeader {
  std::vector<int64_t> dim;
  std::vector<double> vox;
  std::vector<int> layout;        // signed rank per axis: +1 for "+0", -3 for "-2", ...
  std::string datatype = "Float32LE";
  std::vector<std::vector<double>> transform;   // 3 rows of 4
  double scale_offset = 0.0, scale_multiplier = 1.0;
  std::vector<std::pair<std::string, std::string>> keyval;   // other keys, in order

  int64_t nvoxels() const;
  std::string get(const std::string& key) const;   // joined with '\n' when repeated
  void set(const std::string& key, const std::string& value);   // replaces all lines of key
  void erase(const std::string& key);
};

struct MifImage {
  MifHeader h;
  std::vector<float> data;   // values (after intensity scaling), canonical order: axis 0 fastest
};

// Reads any of Float32/Float64/(U)Int8/16/32 in either byte order into float.
MifImage read_mif(const std::string& path);
// Writes Float32LE; `volume_contiguous` puts axis 3 fastest in the file (layout +1,+2,+3,+0),
// otherwise axis 0 fastest.  data is in canonical order.
void write_mif(const std::string& path, const MifHeader& h, const std::vector<float>& data,
               bool volume_contiguous);
// Writes UInt32LE (a fixel index image), axis 0 fastest; data in canonical order.
void write_mif(const std::string& path, const MifHeader& h, const std::vector<uint32_t>& data);
// dst[j][i] = src[i][j] for src [nr][nc] row-major (between an image's volume-slowest canonical
// order and voxel rows), in cache-sized tiles.
void transpose_blocked(const float* src, int64_t nr, int64_t nc, float* dst);

}  // namespace mrcsd
