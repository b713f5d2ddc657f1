// fod2fixel on the GPU: the reference command's arguments, options, checks and fixel directory
// (cmd/fod2fixel.cpp:44-332, with the FMLS options of src/dwi/fmls.cpp:28-91), the per-voxel
// Segmenter run as one batched call of libmrcsd.so (include/mrfixel.h).
//
//   fod2fixel fod fixel_directory [-afd image] [-peak_amp image] [-disp image] [-mask image]
//             [-maxnum n] [-dirpeak] [-fmls_integral value] [-fmls_peak_value value]
//             [-fmls_no_thresholds] [-fmls_lobe_merge_ratio value] [-force] [-quiet] [-info]
//             [-nthreads n]
//
// Options may be shortened to any unambiguous prefix, as the reference's parser allows
// (dwi2response tournier passes -peak for -peak_amp).  The fixels are written in voxel order
// (axis 0 fastest), the order in which FODQueueWriter visits voxels (fmls.h:188-210); the
// reference's receiver takes them in the order its threads finish.  -nthreads is accepted and
// ignored; -nii is refused: this build writes .mif only.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <dirent.h>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mrfixel.h"
#include "host/mif.h"
#include "host/nifti.h"

namespace {

bool g_quiet = false, g_info = false;

void console(const std::string& s) {
  if (!g_quiet) std::fprintf(stderr, "fod2fixel: %s\n", s.c_str());
}
void warn(const std::string& s) { std::fprintf(stderr, "fod2fixel: [WARNING] %s\n", s.c_str()); }
void info(const std::string& s) {
  if (g_info) std::fprintf(stderr, "fod2fixel: [INFO] %s\n", s.c_str());
}

void check(int rc) {
  if (rc != MRCSD_OK) throw std::runtime_error(mrcsd_last_error());
}

// core/math/SH.h:70-73
int lforn(int64_t N) { return N ? 2 * int(std::floor((std::sqrt(float(1 + 8 * N)) - 3.0) / 4.0)) : 0; }

struct OptSpec {
  const char* name;
  int nargs;
};
const OptSpec kOptions[] = {{"afd", 1},           {"peak_amp", 1},        {"disp", 1},
                            {"mask", 1},          {"maxnum", 1},          {"nii", 0},
                            {"dirpeak", 0},       {"fmls_integral", 1},   {"fmls_peak_value", 1},
                            {"fmls_no_thresholds", 0}, {"fmls_lobe_merge_ratio", 1}, {"force", 0},
                            {"quiet", 0},         {"info", 0},            {"debug", 0},
                            {"nthreads", 1}};

// the option a (possibly shortened) name stands for: an exact name, else the only one it starts
const OptSpec& resolve(const std::string& o) {
  const OptSpec* hit = nullptr;
  int n = 0;
  for (const auto& s : kOptions) {
    if (o == s.name) return s;
    if (std::string(s.name).compare(0, o.size(), o) == 0) {
      hit = &s;
      ++n;
    }
  }
  if (n == 1) return *hit;
  if (n == 0) throw std::runtime_error("unknown option \"-" + o + "\"");
  throw std::runtime_error("several matches possible for option \"-" + o + "\"");
}

struct Args {
  std::vector<std::string> pos;
  std::string afd, peak, disp, mask;
  int maxnum = 0;
  bool dirpeak = false, nii = false, no_thresholds = false;
  bool has_integral = false, has_peak_value = false;
  double integral = 0.0, peak_value = 0.1, merge_ratio = 1.0;   // fmls.h:31-33
};

double parse_float(const std::string& opt, const std::string& v, double lo, double hi) {
  size_t used = 0;
  double x = 0.0;
  try {
    x = std::stod(v, &used);
  } catch (const std::exception&) {
    used = 0;
  }
  if (used != v.size()) throw std::runtime_error("value supplied to option \"-" + opt + "\" is not a number: " + v);
  if (x < lo || x > hi)
    throw std::runtime_error("value supplied to option \"-" + opt + "\" is out of bounds (valid range: " +
                             std::to_string(lo) + " to " + std::to_string(hi) + ", value supplied: " + v + ")");
  return x;
}

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1]) && s[1] != '.') {
      const OptSpec& o = resolve(s.substr(s[1] == '-' ? 2 : 1));
      if (i + o.nargs >= argc) throw std::runtime_error("no value supplied for option \"-" + std::string(o.name) + "\"");
      const std::string name = o.name;
      const std::string v = o.nargs ? argv[++i] : "";
      if (name == "afd") a.afd = v;
      else if (name == "peak_amp") a.peak = v;
      else if (name == "disp") a.disp = v;
      else if (name == "mask") a.mask = v;
      else if (name == "maxnum") {
        const double x = parse_float(name, v, 1.0, 2147483647.0);
        if (x != std::floor(x)) throw std::runtime_error("value supplied to option \"-maxnum\" is not an integer: " + v);
        a.maxnum = int(x);
      }
      else if (name == "nii") a.nii = true;
      else if (name == "dirpeak") a.dirpeak = true;
      else if (name == "fmls_integral") { a.has_integral = true; a.integral = parse_float(name, v, 0.0, INFINITY); }
      else if (name == "fmls_peak_value") { a.has_peak_value = true; a.peak_value = parse_float(name, v, 0.0, INFINITY); }
      else if (name == "fmls_no_thresholds") a.no_thresholds = true;
      else if (name == "fmls_lobe_merge_ratio") a.merge_ratio = parse_float(name, v, 0.0, 1.0);
      else if (name == "quiet") g_quiet = true;
      else if (name == "info") g_info = true;
      // force, debug, nthreads: accepted (force cannot apply to a directory, fixel/helpers.h:206-210)
    } else {
      a.pos.push_back(s);
    }
  }
  return a;
}

bool exists(const std::string& p, bool* is_dir) {
  struct stat st;
  if (::stat(p.c_str(), &st) != 0) return false;
  *is_dir = S_ISDIR(st.st_mode);
  return true;
}

bool dir_empty(const std::string& p) {
  DIR* d = ::opendir(p.c_str());
  if (!d) throw std::runtime_error("cannot read directory \"" + p + "\"");
  bool empty = true;
  while (dirent* e = ::readdir(d)) {
    const std::string n = e->d_name;
    if (n != "." && n != "..") {
      empty = false;
      break;
    }
  }
  ::closedir(d);
  return empty;
}

std::string join(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

int run(int argc, char** argv) {
  const auto t0 = std::chrono::steady_clock::now();
  Args a = parse_args(argc, argv);
  if (a.pos.size() != 2) throw std::runtime_error("usage: fod2fixel fod fixel_directory [options]");
  if (a.nii) throw std::runtime_error("-nii: this build writes the fixel directory as .mif images only");
  mrcsd::MifImage fod = mrcsd::read_image(a.pos[0]);
  // Math::SH::check (SH.h:745-751)
  if (fod.h.dim.size() != 4)
    throw std::runtime_error("image \"" + a.pos[0] + "\" does not contain SH coefficients - not 4D");
  const int64_t nsh = fod.h.dim[3];
  const int lmax = lforn(nsh);
  if (lmax % 2 || (lmax + 1) * (lmax + 2) / 2 != nsh)
    throw std::runtime_error("image \"" + a.pos[0] + "\" does not contain SH coefficients - unexpected number of coefficients");
  const int64_t nspatial = fod.h.dim[0] * fod.h.dim[1] * fod.h.dim[2];

  std::vector<uint8_t> mask;
  if (!a.mask.empty()) {
    mrcsd::MifImage m = mrcsd::read_image(a.mask);
    if (m.h.dim.size() < 3 || m.h.dim[0] != fod.h.dim[0] || m.h.dim[1] != fod.h.dim[1] || m.h.dim[2] != fod.h.dim[2])
      throw std::runtime_error("Cannot use image \"" + a.mask + "\" as mask image; dimensions do not match FOD image");
    mask.resize(size_t(nspatial));
    for (int64_t v = 0; v < nspatial; ++v) mask[size_t(v)] = m.data[size_t(v)] != 0.0f;   // Image<float>::value()
  }

  // Fixel::check_fixel_directory (path, create_if_missing, check_if_empty)
  const std::string& dir = a.pos[1];
  bool is_dir = false;
  if (!exists(dir, &is_dir)) {
    if (::mkdir(dir.c_str(), 0777) != 0) throw std::runtime_error("error creating folder \"" + dir + "\"");
  } else if (!is_dir) {
    throw std::runtime_error(dir + " is not a directory");
  } else if (!dir_empty(dir)) {
    throw std::runtime_error("Output fixel directory \"" + dir + "\" is not empty");
  }

  // load_fmls_thresholds (fmls.cpp:56-91)
  double ithr = 0.0, pthr = 0.1;
  if (a.no_thresholds) ithr = pthr = 0.0;
  if (a.has_integral) {
    if (a.no_thresholds) warn("Option -fmls_integral ignored: -fmls_no_thresholds overrides this");
    else ithr = a.integral;
  }
  if (a.has_peak_value) {
    if (a.no_thresholds) warn("Option -fmls_peak_value ignored: -fmls_no_thresholds overrides this");
    else pthr = a.peak_value;
  }

  // voxel rows (coefficient axis contiguous)
  std::vector<float> rows(size_t(nspatial) * size_t(nsh));
  mrcsd::transpose_blocked(fod.data.data(), nsh, nspatial, rows.data());
  const auto t1 = std::chrono::steady_clock::now();

  mrfixel_options opt;
  mrfixel_options_default(&opt);
  opt.n_sh = int(nsh);
  opt.integral_threshold = ithr;
  opt.peak_value_threshold = pthr;
  opt.lobe_merge_ratio = a.merge_ratio;
  opt.maxnum = a.maxnum;
  opt.dir_from_peak = a.dirpeak ? 1 : 0;
  mrfixel_ctx* ctx = nullptr;
  check(mrfixel_create(&opt, &ctx));
  console("segmenting FODs in " + std::to_string(nspatial) + " voxels (GPU)");
  std::vector<uint32_t> count(static_cast<size_t>(nspatial)), offset(static_cast<size_t>(nspatial));
  size_t cap = size_t(4 * nspatial) + 64;
  std::vector<float> fixels;
  uint64_t total = 0;
  const auto t2 = std::chrono::steady_clock::now();
  for (;;) {
    fixels.assign(5 * cap, 0.0f);
    check(mrfixel_segment_host(ctx, rows.data(), int(nsh), mask.empty() ? nullptr : mask.data(), size_t(nspatial),
                               count.data(), offset.data(), fixels.data(), cap, nullptr, &total));
    if (total <= cap) break;
    cap = size_t(total);
  }
  const auto t3 = std::chrono::steady_clock::now();
  mrfixel_destroy(ctx);

  // Segmented_FOD_receiver::commit (fod2fixel.cpp:177-283): nothing is written without fixels
  if (total > 0) {
    const int64_t nf = int64_t(total);
    mrcsd::MifHeader ih = fod.h;
    ih.dim[3] = 2;
    ih.set("nfixels", std::to_string(nf));
    std::vector<uint32_t> index(static_cast<size_t>(2 * nspatial));
    for (int64_t v = 0; v < nspatial; ++v) {
      index[size_t(v)] = count[size_t(v)];
      index[size_t(nspatial + v)] = offset[size_t(v)];
    }
    mrcsd::write_mif(join(dir, "index.mif"), ih, index);
    // fixel data files: the FOD header as n x m x 1, float32
    mrcsd::MifHeader dh = fod.h;
    dh.dim = {nf, 3, 1};
    if (dh.vox.size() > 3) dh.vox.resize(3);
    std::vector<float> d(static_cast<size_t>(3 * nf));
    for (int64_t f = 0; f < nf; ++f)
      for (int c = 0; c < 3; ++c) d[size_t(c * nf + f)] = fixels[size_t(5 * f + c)];
    mrcsd::write_mif(join(dir, "directions.mif"), dh, d, false);
    dh.dim = {nf, 1, 1};
    auto scalar = [&](const std::string& name, int which) {
      if (name.empty()) return;
      std::vector<float> x(static_cast<size_t>(nf));
      for (int64_t f = 0; f < nf; ++f) {
        const float integral = fixels[size_t(5 * f + 3)], peak = fixels[size_t(5 * f + 4)];
        x[size_t(f)] = which == 0 ? integral : which == 1 ? peak : integral / peak;
      }
      mrcsd::write_mif(join(dir, name), dh, x, false);
    };
    scalar(a.afd, 0);
    scalar(a.peak, 1);
    scalar(a.disp, 2);
  }
  const auto t4 = std::chrono::steady_clock::now();
  auto sec = [](auto x, auto y) { return std::chrono::duration<double>(y - x).count(); };
  char buf[220];
  std::snprintf(buf, sizeof buf,
                "%llu fixels; timing: read %.3f s, setup %.3f s, GPU segmentation incl. transfers %.3f s, write %.3f s, "
                "total %.3f s",
                (unsigned long long)total, sec(t0, t1), sec(t1, t2), sec(t2, t3), sec(t3, t4), sec(t0, t4));
  info(buf);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "fod2fixel: [ERROR] %s\n", e.what());
    return 1;
  }
}
