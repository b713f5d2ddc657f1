// sh2peaks on the GPU: the reference command's arguments, options, checks and output
// (cmd/sh2peaks.cpp:30-360), with the per-voxel Processor run as one batched call of
// libmrcsd.so (include/mrpeaks.h).
//
//   sh2peaks SH output [-num n] [-direction phi theta]... [-peaks image] [-threshold value]
//            [-seeds file] [-mask image] [-fast] [-force] [-quiet] [-info] [-nthreads n]
//
// -fast takes the Legendre values from the reference's lookup table (SH.h:402-470); -nthreads is
// accepted and ignored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mrpeaks.h"
#include "host/mif.h"
#include "host/nifti.h"

namespace {

bool g_quiet = false, g_info = false;

void console(const std::string& s) {
  if (!g_quiet) std::fprintf(stderr, "sh2peaks: %s\n", s.c_str());
}
void info(const std::string& s) {
  if (g_info) std::fprintf(stderr, "sh2peaks: [INFO] %s\n", s.c_str());
}

void check(int rc) {
  if (rc != MRCSD_OK) throw std::runtime_error(mrcsd_last_error());
}

bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// a text matrix: rows of numbers separated by spaces, commas or tabs; '#' starts a comment
std::vector<std::vector<double>> load_rows(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open file \"" + path + "\"");
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(f, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    for (char& c : line)
      if (c == ',' || c == ';' || c == '\t' || c == '\r') c = ' ';
    std::istringstream ls(line);
    std::vector<double> r;
    std::string tok;
    while (ls >> tok) r.push_back(std::stod(tok));
    if (!r.empty()) rows.push_back(r);
  }
  if (rows.empty()) throw std::runtime_error("no data in file \"" + path + "\"");
  return rows;
}

// core/math/SH.h:70-73 and :745-751
int lforn(int64_t N) { return N ? 2 * int(std::floor((std::sqrt(float(1 + 8 * N)) - 3.0) / 4.0)) : 0; }

struct Args {
  std::vector<std::string> pos;
  std::string peaks, seeds, mask;
  std::vector<std::pair<float, float>> directions;
  int num = 3;
  bool has_num = false, has_thr = false, force = false, fast = false;
  double thr = -INFINITY;
};

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("no value supplied for option \"" + s + "\"");
      return argv[++i];
    };
    if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1]) && s[1] != '.') {
      const std::string o = s.substr(s[1] == '-' ? 2 : 1);
      if (o == "num") { a.has_num = true; a.num = std::stoi(next()); }
      else if (o == "direction") {
        const float phi = std::stof(next());     // to<float> (sh2peaks.cpp:331)
        const float theta = std::stof(next());
        a.directions.emplace_back(phi, theta);
      }
      else if (o == "peaks") a.peaks = next();
      else if (o == "threshold") { a.has_thr = true; a.thr = std::stod(next()); }
      else if (o == "seeds") a.seeds = next();
      else if (o == "mask") a.mask = next();
      else if (o == "fast") a.fast = true;
      else if (o == "nthreads") next();
      else if (o == "force") a.force = true;
      else if (o == "quiet") g_quiet = true;
      else if (o == "info") g_info = true;
      else throw std::runtime_error("unknown option \"" + s + "\"");
    } else {
      a.pos.push_back(s);
    }
  }
  return a;
}

int run(int argc, char** argv) {
  const auto t0 = std::chrono::steady_clock::now();
  Args a = parse_args(argc, argv);
  if (a.pos.size() != 2) throw std::runtime_error("usage: sh2peaks SH output [options]");
  if (exists(a.pos[1]) && !a.force)
    throw std::runtime_error("output image \"" + a.pos[1] + "\" already exists (use -force option to force overwrite)");
  mrcsd::MifImage sh = mrcsd::read_image(a.pos[0]);
  if (sh.h.dim.size() < 4) throw std::runtime_error("image \"" + a.pos[0] + "\" does not contain SH coefficients - not 4D");
  const int64_t nsh = sh.h.dim[3];
  const int lmax = lforn(nsh);
  if (lmax % 2 || (lmax + 1) * (lmax + 2) / 2 != nsh)
    throw std::runtime_error("image \"" + a.pos[0] + "\" does not contain SH coefficients - unexpected number of coefficients");
  const int64_t nspatial = sh.h.dim[0] * sh.h.dim[1] * sh.h.dim[2];

  std::vector<uint8_t> mask;
  if (!a.mask.empty()) {
    mrcsd::MifImage m = mrcsd::read_image(a.mask);
    if (m.h.dim.size() < 3 || m.h.dim[0] != sh.h.dim[0] || m.h.dim[1] != sh.h.dim[1] || m.h.dim[2] != sh.h.dim[2])
      throw std::runtime_error("dimension mismatch between \"" + a.pos[0] + "\" and mask \"" + a.mask + "\"");
    mask.resize(size_t(nspatial));
    for (int64_t v = 0; v < nspatial; ++v) mask[size_t(v)] = m.data[size_t(v)] != 0.0f;
  }
  std::vector<float> seeds;
  if (!a.seeds.empty()) {
    auto r = load_rows(a.seeds);
    for (auto& row : r) {
      if (row.size() != 2) throw std::runtime_error("expecting 2 columns for search directions matrix");
      seeds.push_back(float(row[0]));
      seeds.push_back(float(row[1]));
    }
  }
  // -direction takes degrees: Math::pi * value / 180.0 in double, then the float Direction
  // (sh2peaks.cpp:328-332)
  std::vector<float> dirs;
  for (auto& d : a.directions) {
    dirs.push_back(float(M_PI * double(d.first) / 180.0));
    dirs.push_back(float(M_PI * double(d.second) / 180.0));
  }
  int npeaks = a.num;
  if (!dirs.empty()) npeaks = int(a.directions.size());
  std::vector<float> ipeaks;
  int ld_ip = 0;
  if (!a.peaks.empty()) {
    if (!dirs.empty())
      throw std::runtime_error("you can't specify both a peaks file and orientations to be estimated at the same time");
    mrcsd::MifImage p = mrcsd::read_image(a.peaks);
    if (p.h.dim.size() < 4 || p.h.dim[0] != sh.h.dim[0] || p.h.dim[1] != sh.h.dim[1] || p.h.dim[2] != sh.h.dim[2])
      throw std::runtime_error("dimension mismatch between \"" + a.pos[0] + "\" and \"" + a.peaks + "\"");
    npeaks = int(p.h.dim[3] / 3);
    ld_ip = int(p.h.dim[3]);
    ipeaks.resize(size_t(nspatial) * size_t(ld_ip));
    mrcsd::transpose_blocked(p.data.data(), ld_ip, nspatial, ipeaks.data());
  }

  // voxel rows (coefficient axis contiguous), every voxel of the image
  std::vector<float> rows(size_t(nspatial) * size_t(nsh));
  mrcsd::transpose_blocked(sh.data.data(), nsh, nspatial, rows.data());
  const auto t1 = std::chrono::steady_clock::now();

  mrpeaks_options opt;
  mrpeaks_options_default(&opt);
  opt.n_sh = int(nsh);
  opt.npeaks = npeaks;
  opt.threshold = float(a.thr);
  opt.fast = a.fast ? 1 : 0;
  if (!seeds.empty()) { opt.seeds = seeds.data(); opt.n_seeds = int(seeds.size() / 2); }
  if (!dirs.empty()) { opt.directions = dirs.data(); opt.n_directions = int(dirs.size() / 2); }
  mrpeaks_ctx* ctx = nullptr;
  check(mrpeaks_create(&opt, &ctx));
  npeaks = mrpeaks_npeaks(ctx);
  console("estimating peak directions in " + std::to_string(nspatial) + " voxels (GPU)");
  std::vector<float> out(size_t(nspatial) * size_t(3 * npeaks));
  const auto t2 = std::chrono::steady_clock::now();
  check(mrpeaks_run_host(ctx, rows.data(), int(nsh), mask.empty() ? nullptr : mask.data(),
                         ipeaks.empty() ? nullptr : ipeaks.data(), ld_ip, size_t(nspatial), out.data()));
  const auto t3 = std::chrono::steady_clock::now();
  mrpeaks_destroy(ctx);

  // output: the SH image's header with 3 npeaks volumes of float32 (sh2peaks.cpp:337-349)
  mrcsd::MifHeader h = sh.h;
  h.dim[3] = 3 * npeaks;
  std::vector<float> data(size_t(nspatial) * size_t(3 * npeaks));
  mrcsd::transpose_blocked(out.data(), nspatial, 3 * npeaks, data.data());
  mrcsd::write_mif(a.pos[1], h, data, true);
  const auto t4 = std::chrono::steady_clock::now();
  auto sec = [](auto x, auto y) { return std::chrono::duration<double>(y - x).count(); };
  char buf[200];
  std::snprintf(buf, sizeof buf, "timing: read %.3f s, setup %.3f s, GPU peaks incl. transfers %.3f s, write %.3f s, total %.3f s",
                sec(t0, t1), sec(t1, t2), sec(t2, t3), sec(t3, t4), sec(t0, t4));
  info(buf);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sh2peaks: [ERROR] %s\n", e.what());
    return 1;
  }
}
