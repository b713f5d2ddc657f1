#include "npz.h"

#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace mrcsd {
namespace npz {
namespace {

uint32_t u16(const unsigned char* p) { return uint32_t(p[0]) | uint32_t(p[1]) << 8; }
uint32_t u32(const unsigned char* p) { return u16(p) | u16(p + 2) << 16; }
uint64_t u64(const unsigned char* p) { return uint64_t(u32(p)) | uint64_t(u32(p + 4)) << 32; }

[[noreturn]] void bad(const std::string& path, const std::string& what) {
  throw std::runtime_error("malformed .npz file \"" + path + "\": " + what);
}

// The .npy payload: magic, version, header dictionary, data.
Array parse_npy(const std::string& path, const std::string& name, const std::vector<unsigned char>& b) {
  static const unsigned char magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
  if (b.size() < 10 || std::memcmp(b.data(), magic, 6) != 0) bad(path, name + " is not an .npy array");
  const int major = b[6];
  size_t hlen, hoff;
  if (major == 1) {
    hlen = u16(&b[8]);
    hoff = 10;
  } else if (major == 2 || major == 3) {
    if (b.size() < 12) bad(path, name + ": truncated header");
    hlen = u32(&b[8]);
    hoff = 12;
  } else {
    bad(path, name + ": unsupported .npy version " + std::to_string(major));
  }
  if (hoff + hlen > b.size()) bad(path, name + ": truncated header");
  const std::string h(reinterpret_cast<const char*>(&b[hoff]), hlen);
  Array a;
  auto value_of = [&](const char* key) -> std::string {
    const size_t k = h.find(std::string("'") + key + "'");
    if (k == std::string::npos) bad(path, name + ": no '" + key + "' in the header");
    size_t v = h.find(':', k);
    if (v == std::string::npos) bad(path, name + ": bad header");
    ++v;
    while (v < h.size() && h[v] == ' ') ++v;
    return h.substr(v);
  };
  {
    const std::string d = value_of("descr");
    if (d.empty() || (d[0] != '\'' && d[0] != '"')) bad(path, name + ": bad descr");
    const size_t e = d.find(d[0], 1);
    if (e == std::string::npos) bad(path, name + ": bad descr");
    a.descr = d.substr(1, e - 1);
  }
  a.fortran_order = value_of("fortran_order").compare(0, 4, "True") == 0;
  {
    const std::string s = value_of("shape");
    if (s.empty() || s[0] != '(') bad(path, name + ": bad shape");
    const size_t e = s.find(')');
    if (e == std::string::npos) bad(path, name + ": bad shape");
    std::string cur;
    for (size_t i = 1; i <= e; ++i) {
      const char c = s[i];
      if (c >= '0' && c <= '9') {
        cur += c;
      } else if (!cur.empty()) {
        a.shape.push_back(size_t(std::stoull(cur)));
        cur.clear();
      }
    }
  }
  size_t item;
  if (a.descr == "<f4" || a.descr == "<i4") item = 4;
  else if (a.descr == "<f8" || a.descr == "<i8") item = 8;
  else if (a.descr == "|u1" || a.descr == "<u1") item = 1;
  else bad(path, name + ": unsupported dtype " + a.descr);
  const size_t n = a.count() * item;
  if (hoff + hlen + n != b.size()) bad(path, name + ": data size does not match its shape");
  a.data.assign(b.begin() + long(hoff + hlen), b.end());
  return a;
}

}  // namespace

size_t Array::count() const {
  size_t n = 1;
  for (size_t s : shape) n *= s;
  return n;
}

std::vector<double> Array::as_double() const {
  const size_t n = count();
  std::vector<double> out(n);
  for (size_t i = 0; i < n; ++i) {
    if (descr == "<f4") {
      float v;
      std::memcpy(&v, &data[4 * i], 4);
      out[i] = v;
    } else if (descr == "<f8") {
      std::memcpy(&out[i], &data[8 * i], 8);
    } else if (descr == "<i4") {
      int32_t v;
      std::memcpy(&v, &data[4 * i], 4);
      out[i] = double(v);
    } else if (descr == "<i8") {
      int64_t v;
      std::memcpy(&v, &data[8 * i], 8);
      out[i] = double(v);
    } else {
      out[i] = double(data[i]);
    }
  }
  return out;
}

// The zip's central directory (with its zip64 extensions) gives each member's sizes, compression
// and local header; a member is stored or deflated.
std::map<std::string, Array> load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open file \"" + path + "\"");
  std::vector<unsigned char> z((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (z.size() < 22) bad(path, "too short");
  // end of central directory record: the last signature in the final 64 KiB + 22 bytes
  size_t eocd = std::string::npos;
  for (size_t i = z.size() - 22 + 1; i-- > 0 && z.size() - i <= 65557;)
    if (u32(&z[i]) == 0x06054b50u) {
      eocd = i;
      break;
    }
  if (eocd == std::string::npos) bad(path, "no end of central directory");
  uint64_t nent = u16(&z[eocd + 10]), cd_off = u32(&z[eocd + 16]);
  if ((nent == 0xFFFF || cd_off == 0xFFFFFFFFu) && eocd >= 20 && u32(&z[eocd - 20]) == 0x07064b50u) {
    const uint64_t z64 = u64(&z[eocd - 20 + 8]);
    if (z64 + 56 > z.size() || u32(&z[z64]) != 0x06064b50u) bad(path, "bad zip64 end of central directory");
    nent = u64(&z[z64 + 32]);
    cd_off = u64(&z[z64 + 48]);
  }
  std::map<std::string, Array> out;
  size_t p = size_t(cd_off);
  for (uint64_t e = 0; e < nent; ++e) {
    if (p + 46 > z.size() || u32(&z[p]) != 0x02014b50u) bad(path, "bad central directory entry");
    const uint32_t method = u16(&z[p + 10]);
    uint64_t csize = u32(&z[p + 20]), usize = u32(&z[p + 24]);
    const uint32_t nlen = u16(&z[p + 28]), xlen = u16(&z[p + 30]), clen = u16(&z[p + 32]);
    uint64_t loff = u32(&z[p + 42]);
    if (p + 46 + nlen + xlen > z.size()) bad(path, "truncated central directory");
    std::string name(reinterpret_cast<const char*>(&z[p + 46]), nlen);
    // zip64 extra field: the 0xFFFFFFFF fields in order (uncompressed, compressed, offset)
    for (size_t x = p + 46 + nlen; x + 4 <= p + 46 + nlen + xlen;) {
      const uint32_t id = u16(&z[x]), len = u16(&z[x + 2]);
      if (id == 0x0001) {
        size_t q = x + 4;
        if (usize == 0xFFFFFFFFu && q + 8 <= x + 4 + len) { usize = u64(&z[q]); q += 8; }
        if (csize == 0xFFFFFFFFu && q + 8 <= x + 4 + len) { csize = u64(&z[q]); q += 8; }
        if (loff == 0xFFFFFFFFu && q + 8 <= x + 4 + len) { loff = u64(&z[q]); }
      }
      x += 4 + len;
    }
    p += 46 + nlen + xlen + clen;
    if (loff + 30 > z.size() || u32(&z[loff]) != 0x04034b50u) bad(path, "bad local header of " + name);
    const size_t data = size_t(loff) + 30 + u16(&z[loff + 26]) + u16(&z[loff + 28]);
    if (data + csize > z.size()) bad(path, "truncated member " + name);
    std::vector<unsigned char> body;
    if (method == 0) {
      if (csize != usize) bad(path, "stored member " + name + " with differing sizes");
      body.assign(z.begin() + long(data), z.begin() + long(data + csize));
    } else if (method == 8) {
      body.resize(size_t(usize));
      z_stream s{};
      if (inflateInit2(&s, -15) != Z_OK) bad(path, "zlib");
      s.next_in = &z[data];
      s.avail_in = uInt(csize);
      s.next_out = body.data();
      s.avail_out = uInt(usize);
      const int rc = inflate(&s, Z_FINISH);
      inflateEnd(&s);
      if (rc != Z_STREAM_END || s.total_out != usize) bad(path, "cannot inflate member " + name);
    } else {
      bad(path, "member " + name + " uses an unsupported compression method");
    }
    if (name.size() > 4 && name.compare(name.size() - 4, 4, ".npy") == 0) name.resize(name.size() - 4);
    out[name] = parse_npy(path, name, body);
  }
  return out;
}

}  // namespace npz
}  // namespace mrcsd
