#include "mif.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace mrcsd {

namespace {

std::string trim(const std::string& s) {
  const auto a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  const auto b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string& s, const std::string& seps) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (seps.find(c) != std::string::npos) {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

std::string lower(std::string s) {
  for (char& c : s) c = char(std::tolower(c));
  return s;
}

struct DType {
  int bytes = 4;
  bool is_float = true, is_signed = true, big_endian = false;
};

DType parse_dtype(const std::string& spec) {
  std::string s = lower(spec);
  DType d;
  d.big_endian = s.size() > 2 && s.substr(s.size() - 2) == "be";
  if (s.size() > 2 && (s.substr(s.size() - 2) == "le" || s.substr(s.size() - 2) == "be")) s = s.substr(0, s.size() - 2);
  if (s == "float32") { d.bytes = 4; d.is_float = true; }
  else if (s == "float64") { d.bytes = 8; d.is_float = true; }
  else if (s == "int8") { d.bytes = 1; d.is_float = false; }
  else if (s == "uint8") { d.bytes = 1; d.is_float = false; d.is_signed = false; }
  else if (s == "int16") { d.bytes = 2; d.is_float = false; }
  else if (s == "uint16") { d.bytes = 2; d.is_float = false; d.is_signed = false; }
  else if (s == "int32") { d.bytes = 4; d.is_float = false; }
  else if (s == "uint32") { d.bytes = 4; d.is_float = false; d.is_signed = false; }
  else if (s == "bit") throw std::runtime_error("bitwise .mif data is not supported: " + spec);
  else throw std::runtime_error("unsupported datatype \"" + spec + "\"");
  return d;
}

double decode(const unsigned char* p, const DType& d) {
  unsigned char b[8];
  for (int i = 0; i < d.bytes; ++i) b[i] = d.big_endian ? p[d.bytes - 1 - i] : p[i];
  if (d.is_float) {
    if (d.bytes == 4) { float f; std::memcpy(&f, b, 4); return f; }
    double v; std::memcpy(&v, b, 8); return v;
  }
  switch (d.bytes) {
    case 1: return d.is_signed ? double(int8_t(b[0])) : double(b[0]);
    case 2: { uint16_t u; std::memcpy(&u, b, 2); return d.is_signed ? double(int16_t(u)) : double(u); }
    default: { uint32_t u; std::memcpy(&u, b, 4); return d.is_signed ? double(int32_t(u)) : double(u); }
  }
}

// Byte strides of each axis in the file for a layout of signed ranks (sign ignored: canonical
// order is always returned, negative strides flip the axis in the file).
std::vector<int64_t> file_strides(const MifHeader& h, std::vector<bool>& flip) {
  const size_t nd = h.dim.size();
  std::vector<size_t> order(nd);
  for (size_t i = 0; i < nd; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return std::abs(h.layout[a]) < std::abs(h.layout[b]); });
  std::vector<int64_t> st(nd);
  int64_t s = 1;
  flip.assign(nd, false);
  for (size_t k = 0; k < nd; ++k) {
    st[order[k]] = s;
    s *= h.dim[order[k]];
    flip[order[k]] = h.layout[order[k]] < 0;
  }
  return st;
}

// Runs body(r0, r1) over [0, nrows) split into contiguous ranges on up to 8 host threads (one
// range when the image is small).
template <class F>
void parallel_rows(int64_t nrows, F&& body) {
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({8, int64_t(std::thread::hardware_concurrency()), nrows / 256}));
  if (nt <= 1) { body(0, nrows); return; }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] { body(nrows * t / nt, nrows * (t + 1) / nt); });
  for (auto& x : th) x.join();
}

// Calls row(f, c) for every row along axis 0, with f the file offset (in elements) of the row's
// element x = 0 and c its canonical offset (axis 0 fastest).  Rows are visited with the other axes
// nested in the order of their file strides (the smallest innermost), so consecutive rows share
// the cache lines of an interleaved file layout; ranges of rows go to different host threads.
template <class F>
void for_each_row(const std::vector<int64_t>& dim, const std::vector<int64_t>& st, const std::vector<bool>& flip,
                  F&& row) {
  const size_t nd = dim.size();
  std::vector<size_t> axes;
  for (size_t a = 1; a < nd; ++a) axes.push_back(a);
  std::sort(axes.begin(), axes.end(), [&](size_t a, size_t b) { return st[a] < st[b]; });
  std::vector<int64_t> cst(nd, 1);
  for (size_t a = 1; a < nd; ++a) cst[a] = cst[a - 1] * dim[a - 1];
  int64_t nrows = 1;
  for (size_t a : axes) nrows *= dim[a];
  const int64_t f0 = flip[0] ? (dim[0] - 1) * st[0] : 0;
  parallel_rows(nrows, [&](int64_t r0, int64_t r1) {
    std::vector<int64_t> idx(nd, 0);
    int64_t rem = r0;
    for (size_t a : axes) { idx[a] = rem % dim[a]; rem /= dim[a]; }
    for (int64_t r = r0; r < r1; ++r) {
      int64_t f = f0, c = 0;
      for (size_t a : axes) {
        f += (flip[a] ? dim[a] - 1 - idx[a] : idx[a]) * st[a];
        c += idx[a] * cst[a];
      }
      row(f, c);
      for (size_t a : axes) {
        if (++idx[a] < dim[a]) break;
        idx[a] = 0;
      }
    }
  });
}

}  // namespace

int64_t MifHeader::nvoxels() const {
  int64_t n = 1;
  for (auto d : dim) n *= d;
  return n;
}

std::string MifHeader::get(const std::string& key) const {
  std::string out;
  for (const auto& kv : keyval)
    if (kv.first == key) out += (out.empty() ? "" : "\n") + kv.second;
  return out;
}

void MifHeader::erase(const std::string& key) {
  keyval.erase(std::remove_if(keyval.begin(), keyval.end(), [&](const auto& kv) { return kv.first == key; }),
               keyval.end());
}

void MifHeader::set(const std::string& key, const std::string& value) {
  erase(key);
  std::istringstream in(value);
  std::string line;
  while (std::getline(in, line)) keyval.emplace_back(key, line);
}

MifImage read_mif(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("failed to open image \"" + path + "\"");
  std::string line;
  std::getline(in, line);
  if (trim(line) != "mrtrix image") throw std::runtime_error("image \"" + path + "\" is not in MRtrix format");
  MifImage img;
  MifHeader& h = img.h;
  std::string file_spec;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line == "END") break;
    const auto colon = line.find(':');
    if (colon == std::string::npos) continue;
    const std::string key = trim(line.substr(0, colon)), value = trim(line.substr(colon + 1));
    const std::string lk = lower(key);
    if (lk == "dim") { for (auto& t : split(value, ",")) h.dim.push_back(std::stoll(t)); }
    else if (lk == "vox") { for (auto& t : split(value, ",")) h.vox.push_back(std::stod(t)); }
    else if (lk == "layout") {
      for (auto& t : split(value, ",")) {
        const std::string u = trim(t);
        const int sign = u[0] == '-' ? -1 : 1;
        const int rank = std::stoi(u[0] == '+' || u[0] == '-' ? u.substr(1) : u);
        h.layout.push_back(sign * (rank + 1));
      }
    }
    else if (lk == "datatype") h.datatype = value;
    else if (lk == "transform") {
      std::vector<double> row;
      for (auto& t : split(value, ", ")) row.push_back(std::stod(t));
      h.transform.push_back(row);
    }
    else if (lk == "scaling") {
      auto v = split(value, ",");
      if (v.size() != 2) throw std::runtime_error("invalid \"scaling\" specification for MRtrix image \"" + path + "\"");
      h.scale_offset = std::stod(v[0]);
      h.scale_multiplier = std::stod(v[1]);
    }
    else if (lk == "file") file_spec = value;
    else if (!key.empty() && !value.empty()) h.keyval.emplace_back(key, value);
  }
  if (h.dim.empty()) throw std::runtime_error("missing \"dim\" specification for MRtrix image \"" + path + "\"");
  for (auto d : h.dim)
    if (d < 1) throw std::runtime_error("invalid dimensions for MRtrix image \"" + path + "\"");
  if (h.vox.empty()) throw std::runtime_error("missing \"vox\" specification for MRtrix image \"" + path + "\"");
  if (h.layout.size() != h.dim.size())
    throw std::runtime_error("missing \"layout\" specification for MRtrix image \"" + path + "\"");
  if (file_spec.empty()) throw std::runtime_error("missing \"file\" specification for MRtrix image \"" + path + "\"");
  auto fs = split(file_spec, " ");
  std::string data_path = fs[0] == "." ? path : fs[0];
  if (fs[0] != "." && data_path.find('/') == std::string::npos) {
    const auto slash = path.rfind('/');
    if (slash != std::string::npos) data_path = path.substr(0, slash + 1) + data_path;
  }
  const int64_t offset = fs.size() > 1 ? std::stoll(fs[1]) : 0;
  const DType dt = parse_dtype(h.datatype);
  const int64_t n = h.nvoxels();
  std::ifstream din(data_path, std::ios::binary);
  if (!din) throw std::runtime_error("failed to open data file \"" + data_path + "\"");
  din.seekg(offset);
  std::vector<unsigned char> raw(size_t(n) * size_t(dt.bytes));
  din.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size()));
  if (din.gcount() != std::streamsize(raw.size()))
    throw std::runtime_error("image file \"" + data_path + "\" is too small for its header");
  std::vector<bool> flip;
  const auto st = file_strides(h, flip);
  img.data.resize(size_t(n));
  const int64_t nx = h.dim[0];
  const int64_t sx = flip[0] ? -st[0] : st[0];
  const double so = h.scale_offset, sm = h.scale_multiplier;
  const bool plain = dt.is_float && dt.bytes == 4 && !dt.big_endian && so == 0.0 && sm == 1.0;
  for_each_row(h.dim, st, flip, [&](int64_t f, int64_t c) {
    float* dst = &img.data[size_t(c)];
    if (plain) {
      for (int64_t x = 0; x < nx; ++x, f += sx) std::memcpy(&dst[x], &raw[size_t(f) * 4], 4);
    } else {
      for (int64_t x = 0; x < nx; ++x, f += sx) dst[x] = float(so + sm * decode(&raw[size_t(f) * dt.bytes], dt));
    }
  });
  h.datatype = "Float32LE";
  h.scale_offset = 0.0;
  h.scale_multiplier = 1.0;
  return img;
}

namespace {

// T is float (Float32LE) or uint32_t (UInt32LE), both 4 bytes
template <class T>
void write_mif_t(const std::string& path, const MifHeader& hin, const std::vector<T>& data, bool volume_contiguous,
                 const char* dtype) {
  static_assert(sizeof(T) == 4, "4-byte element types only");
  MifHeader h = hin;
  const size_t nd = h.dim.size();
  if (int64_t(data.size()) != h.nvoxels()) throw std::runtime_error("write_mif: data size does not match dimensions");
  h.layout.assign(nd, 0);
  for (size_t a = 0; a < nd; ++a) h.layout[a] = int(a) + 1;
  if (volume_contiguous && nd == 4) h.layout = {2, 3, 4, 1};
  std::ostringstream hdr;
  hdr.precision(17);
  hdr << "mrtrix image\n";
  hdr << "dim: ";
  for (size_t a = 0; a < nd; ++a) hdr << (a ? "," : "") << h.dim[a];
  hdr << "\nvox: ";
  for (size_t a = 0; a < nd; ++a) hdr << (a ? "," : "") << (a < h.vox.size() ? h.vox[a] : 1.0);
  hdr << "\nlayout: ";
  for (size_t a = 0; a < nd; ++a) hdr << (a ? "," : "") << (h.layout[a] > 0 ? "+" : "-") << std::abs(h.layout[a]) - 1;
  hdr << "\ndatatype: " << dtype;
  if (h.transform.size() >= 3) {
    for (int r = 0; r < 3; ++r) {
      hdr << "\ntransform: ";
      for (int c = 0; c < 4; ++c) hdr << (c ? ", " : "") << h.transform[size_t(r)][size_t(c)];
    }
  } else {
    hdr << "\ntransform: 1, 0, 0, 0\ntransform: 0, 1, 0, 0\ntransform: 0, 0, 1, 0";
  }
  for (const auto& kv : h.keyval) hdr << "\n" << kv.first << ": " << kv.second;
  hdr << "\nfile: ";
  std::string head = hdr.str();
  int64_t offset = int64_t(head.size()) + 18;
  offset += (4 - offset % 4) % 4;
  head += ". " + std::to_string(offset) + "\nEND\n";
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("failed to create image \"" + path + "\"");
  out << head;
  const int64_t pad = offset - int64_t(head.size());
  for (int64_t i = 0; i < pad; ++i) out.put('\0');
  std::vector<bool> flip;
  const auto st = file_strides(h, flip);
  std::vector<T> buf(data.size());
  const int64_t nx = h.dim[0];
  for_each_row(h.dim, st, flip, [&](int64_t f, int64_t c) {
    const T* src = &data[size_t(c)];
    if (st[0] == 1) {
      std::memcpy(&buf[size_t(f)], src, size_t(nx) * sizeof(T));
    } else {
      for (int64_t x = 0; x < nx; ++x, f += st[0]) buf[size_t(f)] = src[x];
    }
  });
  out.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size() * sizeof(T)));
  if (!out) throw std::runtime_error("failed writing image \"" + path + "\"");
}

}  // namespace

void write_mif(const std::string& path, const MifHeader& h, const std::vector<float>& data, bool volume_contiguous) {
  write_mif_t(path, h, data, volume_contiguous, "Float32LE");
}

void write_mif(const std::string& path, const MifHeader& h, const std::vector<uint32_t>& data) {
  write_mif_t(path, h, data, false, "UInt32LE");
}

void transpose_blocked(const float* src, int64_t nr, int64_t nc, float* dst) {
  constexpr int64_t B = 64;
  for (int64_t i0 = 0; i0 < nr; i0 += B)
    for (int64_t j0 = 0; j0 < nc; j0 += B) {
      const int64_t i1 = std::min(nr, i0 + B), j1 = std::min(nc, j0 + B);
      for (int64_t i = i0; i < i1; ++i)
        for (int64_t j = j0; j < j1; ++j) dst[size_t(j * nr + i)] = src[size_t(i * nc + j)];
    }
}

}  // namespace mrcsd
