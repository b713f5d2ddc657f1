#include "nifti.h"

#include <zlib.h>

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace mrcsd {
namespace {

bool ends_with(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

// the whole file, decompressed when it is gzip (gzread passes plain files through)
std::vector<unsigned char> slurp(const std::string& path) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("failed to open image \"" + path + "\"");
  std::vector<unsigned char> buf;
  unsigned char chunk[1 << 16];
  for (;;) {
    const int n = gzread(f, chunk, sizeof chunk);
    if (n < 0) {
      gzclose(f);
      throw std::runtime_error("error reading \"" + path + "\"");
    }
    if (n == 0) break;
    buf.insert(buf.end(), chunk, chunk + n);
  }
  gzclose(f);
  return buf;
}

struct Reader {
  const unsigned char* p;
  bool swap;
  template <class T>
  T get(size_t off) const {
    T v;
    unsigned char b[sizeof(T)];
    std::memcpy(b, p + off, sizeof(T));
    if (swap)
      for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
    std::memcpy(&v, b, sizeof(T));
    return v;
  }
};

template <class T>
void convert(const Reader& r, size_t off, size_t n, double slope, double inter, std::vector<float>& out) {
  out.resize(n);
  for (size_t i = 0; i < n; ++i) out[i] = float(slope * double(r.get<T>(off + i * sizeof(T))) + inter);
}

}  // namespace

MifImage read_nifti(const std::string& path) {
  const std::vector<unsigned char> buf = slurp(path);
  if (buf.size() < 348) throw std::runtime_error("image \"" + path + "\" is too small to be NIfTI");
  // sizeof_hdr gives the version and the byte order: 348 for NIfTI-1, 540 for NIfTI-2
  Reader r{buf.data(), false};
  const int32_t native = r.get<int32_t>(0);
  r.swap = true;
  const int32_t swapped = r.get<int32_t>(0);
  r.swap = native != 348;
  if (native == 540 || swapped == 540)
    throw std::runtime_error("NIfTI-2 image \"" + path + "\" is not supported (NIfTI-1 only)");
  if (native != 348 && swapped != 348) throw std::runtime_error("image \"" + path + "\" is not in NIfTI-1 format");
  if (std::memcmp(buf.data() + 344, "n+1", 4) != 0 && std::memcmp(buf.data() + 344, "ni1", 4) != 0)
    throw std::runtime_error("image \"" + path + "\" is not a single-file NIfTI-1 image (magic \"n+1\")");

  MifImage img;
  MifHeader& h = img.h;
  const int ndim = r.get<int16_t>(40);
  if (ndim < 1 || ndim > 7) throw std::runtime_error("image \"" + path + "\": invalid number of dimensions");
  int64_t nvox = 1;
  for (int i = 0; i < ndim; ++i) {
    const int64_t d = r.get<int16_t>(42 + 2 * i);
    if (d < 1) throw std::runtime_error("image \"" + path + "\": invalid dimension");
    h.dim.push_back(d);
    nvox *= d;
  }
  // pixdim[1..ndim], magnitudes (nifti_utils.cpp:198-206)
  double pixdim[8];
  for (int i = 0; i < 8; ++i) pixdim[i] = r.get<float>(76 + 4 * i);
  for (int i = 0; i < ndim; ++i) h.vox.push_back(std::fabs(pixdim[i + 1]));
  // intensity scaling (nifti_utils.cpp:209-217): a zero or non-finite slope means none
  double slope = r.get<float>(112), inter = r.get<float>(116);
  if (!std::isfinite(slope) || slope == 0.0) {
    slope = 1.0;
    inter = 0.0;
  }
  if (!std::isfinite(inter)) inter = 0.0;
  const int16_t datatype = r.get<int16_t>(70);
  const float vox_offset = r.get<float>(108);
  const size_t off = size_t(vox_offset);
  size_t elem = 0;
  switch (datatype) {
    case 2: case 256: elem = 1; break;
    case 4: case 512: elem = 2; break;
    case 8: case 768: case 16: elem = 4; break;
    case 64: case 1024: case 1280: elem = 8; break;
    default: throw std::runtime_error("image \"" + path + "\": unsupported NIfTI data type " + std::to_string(datatype));
  }
  if (off < 348 || off + size_t(nvox) * elem > buf.size())
    throw std::runtime_error("image \"" + path + "\": data past the end of the file");
  switch (datatype) {
    case 2: convert<uint8_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 256: convert<int8_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 4: convert<int16_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 512: convert<uint16_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 8: convert<int32_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 768: convert<uint32_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 16: convert<float>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 64: convert<double>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 1024: convert<int64_t>(r, off, size_t(nvox), slope, inter, img.data); break;
    case 1280: convert<uint64_t>(r, off, size_t(nvox), slope, inter, img.data); break;
  }

  // transform (nifti_utils.cpp:240-330): the s-form when set, else the q-form
  double M[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
  const int16_t qform_code = r.get<int16_t>(252), sform_code = r.get<int16_t>(254);
  if (sform_code) {
    for (int row = 0; row < 3; ++row)
      for (int c = 0; c < 4; ++c) M[row][c] = r.get<float>(280 + 16 * row + 4 * c);
    bool rescale = false;
    for (int a = 0; a < 3 && a < ndim; ++a) {
      const double len = std::sqrt(M[0][a] * M[0][a] + M[1][a] * M[1][a] + M[2][a] * M[2][a]);
      if (std::fabs(pixdim[a + 1] / len - 1.0) > 1e-5) rescale = true;
    }
    for (int a = 0; a < 3 && a < ndim; ++a) {
      const double len = std::sqrt(M[0][a] * M[0][a] + M[1][a] * M[1][a] + M[2][a] * M[2][a]);
      for (int row = 0; row < 3; ++row) M[row][a] /= len;
      h.vox[size_t(a)] = rescale ? len : std::fabs(pixdim[a + 1]);
    }
  } else if (qform_code) {
    double b = r.get<float>(256), c = r.get<float>(260), d = r.get<float>(264);
    double w2 = 1.0 - (b * b + c * c + d * d), a;
    if (w2 < 1.0e-7) {   // normalise the vector part
      const double n = std::sqrt(b * b + c * c + d * d);
      b /= n;
      c /= n;
      d /= n;
      a = 0.0;
    } else {
      a = std::sqrt(w2);
    }
    // the unit quaternion (a, b, c, d) as a rotation matrix (the standard formula)
    const double tx = 2 * b, ty = 2 * c, tz = 2 * d;
    const double twx = tx * a, twy = ty * a, twz = tz * a, txx = tx * b, txy = ty * b, txz = tz * b;
    const double tyy = ty * c, tyz = tz * c, tzz = tz * d;
    const double R[3][3] = {{1 - (tyy + tzz), txy - twz, txz + twy},
                            {txy + twz, 1 - (txx + tzz), tyz - twx},
                            {txz - twy, tyz + twx, 1 - (txx + tyy)}};
    const double qfac = pixdim[0] >= 0.0 ? 1.0 : -1.0;
    for (int row = 0; row < 3; ++row) {
      for (int col = 0; col < 3; ++col) M[row][col] = R[row][col] * (col == 2 ? qfac : 1.0);
    }
    M[0][3] = r.get<float>(268);
    M[1][3] = r.get<float>(272);
    M[2][3] = r.get<float>(276);
  }
  h.transform.assign(3, std::vector<double>(4));
  for (int row = 0; row < 3; ++row)
    for (int c = 0; c < 4; ++c) h.transform[size_t(row)][size_t(c)] = M[row][c];
  h.layout.clear();
  for (int i = 0; i < ndim; ++i) h.layout.push_back(i + 1);   // axis 0 fastest, as stored
  return img;
}

MifImage read_image(const std::string& path) {
  if (ends_with(path, ".nii") || ends_with(path, ".nii.gz")) return read_nifti(path);
  return read_mif(path);
}

}  // namespace mrcsd
