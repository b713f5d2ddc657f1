// amp2response on the GPU: the reference command's arguments, options, checks and output
// (cmd/amp2response.cpp:46-380), with the per-voxel Accumulator and the per-shell solve run by
// libmrcsd.so (include/mrresp.h).
//
//   amp2response amps mask directions response [-isotropic] [-noconstraint] [-directions file]
//                [-shells b1,b2,...] [-lmax l1,l2,...] [-force] [-quiet] [-info] [-nthreads n]
//
// The fibre frame's x and y axes, which the reference draws at random (amp2response.cpp:72-90),
// do not enter the zonal fit; -nthreads is accepted and ignored.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mrresp.h"
#include "host/dense.h"
#include "host/gradient.h"
#include "host/mif.h"
#include "host/nifti.h"

namespace {

bool g_quiet = false, g_info = false;

void console(const std::string& s) {
  if (!g_quiet) std::fprintf(stderr, "amp2response: %s\n", s.c_str());
}
void info(const std::string& s) {
  if (g_info) std::fprintf(stderr, "amp2response: [INFO] %s\n", s.c_str());
}
void check(int rc) {
  if (rc != MRCSD_OK) throw std::runtime_error(mrcsd_last_error());
}
void hip_check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e));
}
bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

std::vector<std::vector<double>> parse_rows(std::istream& in) {
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(in, line)) {
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
  return rows;
}
std::vector<std::vector<double>> load_rows(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open file \"" + path + "\"");
  auto r = parse_rows(f);
  if (r.empty()) throw std::runtime_error("no data in file \"" + path + "\"");
  return r;
}
template <class T>
std::vector<T> parse_list(const std::string& s) {
  std::vector<T> out;
  std::string cur;
  for (char c : s + ",") {
    if (c == ',') {
      if (!cur.empty()) out.push_back(T(std::stod(cur)));
      cur.clear();
    } else if (c != ' ') {
      cur += c;
    }
  }
  return out;
}

mrcsd::Matrix to_matrix(const std::vector<std::vector<double>>& r, int cols, const std::string& what) {
  mrcsd::Matrix m(int(r.size()), cols);
  for (size_t i = 0; i < r.size(); ++i) {
    if (int(r[i].size()) != cols) throw std::runtime_error(what + ": expected " + std::to_string(cols) + " columns");
    for (int c = 0; c < cols; ++c) m(int(i), c) = r[i][size_t(c)];
  }
  return m;
}

struct Args {
  std::vector<std::string> pos;
  std::string directions, lmax, shells;
  bool isotropic = false, noconstraint = false, force = false;
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
      if (o == "isotropic") a.isotropic = true;
      else if (o == "noconstraint") a.noconstraint = true;
      else if (o == "directions") a.directions = next();
      else if (o == "lmax") a.lmax = next();
      else if (o == "shells") a.shells = next();
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
  if (a.pos.size() != 4) throw std::runtime_error("usage: amp2response amps mask directions response [options]");
  if (exists(a.pos[3]) && !a.force)
    throw std::runtime_error("output file \"" + a.pos[3] + "\" already exists (use -force option to force overwrite)");
  mrcsd::MifImage img = mrcsd::read_image(a.pos[0]);
  const int64_t nspatial = img.h.dim.size() >= 3 ? img.h.dim[0] * img.h.dim[1] * img.h.dim[2] : 0;
  const int64_t nvol_img = img.h.dim.size() >= 4 ? img.h.dim[3] : 1;

  // the sampling directions: -directions, the header's "directions", or the DW scheme's shells
  // (amp2response.cpp:207-239)
  std::vector<mrcsd::Matrix> dirs_azel;
  std::vector<std::vector<size_t>> volumes;
  std::vector<MR::DWI::Shell> shell_list;
  bool have_shells = false;
  auto all_volumes = [](size_t n) {
    std::vector<size_t> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = i;
    return v;
  };
  const std::string hdr_dirs = img.h.get("directions");
  if (!a.directions.empty()) {
    dirs_azel.push_back(to_matrix(load_rows(a.directions), 2, "directions file"));
    volumes.push_back(all_volumes(size_t(dirs_azel.back().rows)));
  } else if (!hdr_dirs.empty()) {
    std::istringstream in(hdr_dirs);
    std::vector<double> flat;
    for (auto& r : parse_rows(in)) flat.insert(flat.end(), r.begin(), r.end());
    mrcsd::Matrix d(int(flat.size() / 2), 2);
    for (size_t i = 0; i + 1 < flat.size(); i += 2) {
      d(int(i / 2), 0) = flat[i];
      d(int(i / 2), 1) = flat[i + 1];
    }
    dirs_azel.push_back(d);
    volumes.push_back(all_volumes(size_t(d.rows)));
  } else {
    std::istringstream in(img.h.get("dw_scheme"));
    auto rows = parse_rows(in);
    if (rows.empty()) throw std::runtime_error("no diffusion gradient table found for image \"" + a.pos[0] + "\"");
    const mrcsd::Matrix grad = MR::DWI::get_DW_scheme(to_matrix(rows, 4, "dw_scheme"), int(nvol_img));
    MR::DWI::Shells shells(grad);
    shells.select_shells(false, false, false, parse_list<double>(a.shells));
    have_shells = true;
    for (size_t i = 0; i < shells.count(); ++i) {
      shell_list.push_back(shells[i]);
      volumes.push_back(shells[i].get_volumes());
      dirs_azel.push_back(MR::DWI::gen_direction_matrix(grad, volumes.back()));
    }
  }

  // lmax per shell (amp2response.cpp:241-279)
  std::vector<int> lmax;
  int max_lmax = 0;
  if (a.isotropic) {
    lmax.assign(dirs_azel.size(), 0);
  } else if (!a.lmax.empty()) {
    lmax = parse_list<int>(a.lmax);
    if (lmax.size() != dirs_azel.size())
      throw std::runtime_error("Number of lmax's specified (" + std::to_string(lmax.size()) +
                               ") does not match number of b-value shells (" + std::to_string(dirs_azel.size()) + ")");
    for (int l : lmax) {
      if (l % 2) throw std::runtime_error("Values specified for lmax must be even");
      max_lmax = std::max(max_lmax, l);
    }
    if (have_shells && shell_list[0].is_bzero() && lmax.front()) {
      console("[WARNING] Non-zero lmax requested for " +
              (shell_list[0].get_mean() ? "first shell (mean b=" + std::to_string(shell_list[0].get_mean()) +
                                               "), which MRtrix3 has classified as b=0;"
                                        : std::string("b=0 shell;")));
      console("[WARNING]   unless intended, this is likely to fail, as b=0 contains no orientation contrast");
    }
  } else {
    for (size_t i = 0; i < dirs_azel.size(); ++i) lmax.push_back(!i && have_shells && shell_list.front().is_bzero() ? 0 : 10);
    max_lmax = (have_shells && shell_list.front().is_bzero() && lmax.size() == 1) ? 0 : 10;
  }
  for (int l : lmax)
    if (l < 0 || l > 12) throw std::runtime_error("lmax must be in [0, 12] on the GPU");

  // mask and fibre directions (amp2response.cpp:281-300)
  mrcsd::MifImage mask_img = mrcsd::read_image(a.pos[1]);
  for (int d = 0; d < 3; ++d)
    if (mask_img.h.dim.size() < 3 || mask_img.h.dim[size_t(d)] != img.h.dim[size_t(d)])
      throw std::runtime_error("dimension mismatch between \"" + a.pos[0] + "\" and \"" + a.pos[1] + "\"");
  if (!(mask_img.h.dim.size() == 3 || (mask_img.h.dim.size() == 4 && mask_img.h.dim[3] == 1)))
    throw std::runtime_error("input mask must be a 3D image");
  mrcsd::MifImage dir_img = mrcsd::read_image(a.pos[2]);
  if (dir_img.h.dim.size() < 4 || dir_img.h.dim[3] < 3)
    throw std::runtime_error("input direction image \"" + a.pos[2] + "\" does not have expected dimensions");
  for (int d = 0; d < 3; ++d)
    if (dir_img.h.dim[size_t(d)] != img.h.dim[size_t(d)])
      throw std::runtime_error("dimension mismatch between \"" + a.pos[0] + "\" and \"" + a.pos[2] + "\"");
  std::vector<uint8_t> mask(static_cast<size_t>(nspatial));
  size_t num_voxels = 0;
  for (int64_t v = 0; v < nspatial; ++v) num_voxels += (mask[size_t(v)] = mask_img.data[size_t(v)] != 0.0f);
  if (!num_voxels) throw std::runtime_error("input mask does not contain any voxels");
  console(std::string("estimating response function using ") + (a.noconstraint ? "ordinary" : "constrained") +
          " least-squares from " + std::to_string(num_voxels) + " voxels");

  // voxel rows (volume axis contiguous) on the device
  std::vector<float> amp(size_t(nspatial) * size_t(nvol_img)), fib(size_t(nspatial) * 3);
  mrcsd::transpose_blocked(img.data.data(), nvol_img, nspatial, amp.data());
  mrcsd::transpose_blocked(dir_img.data.data(), 3, nspatial, fib.data());
  const auto t1 = std::chrono::steady_clock::now();
  float *d_amp = nullptr, *d_fib = nullptr;
  uint8_t* d_mask = nullptr;
  hip_check(hipMalloc(&d_amp, amp.size() * sizeof(float)));
  hip_check(hipMalloc(&d_fib, fib.size() * sizeof(float)));
  hip_check(hipMalloc(&d_mask, mask.size()));
  hip_check(hipMemcpy(d_amp, amp.data(), amp.size() * sizeof(float), hipMemcpyHostToDevice));
  hip_check(hipMemcpy(d_fib, fib.data(), fib.size() * sizeof(float), hipMemcpyHostToDevice));
  hip_check(hipMemcpy(d_mask, mask.data(), mask.size(), hipMemcpyHostToDevice));

  const int nz_max = max_lmax / 2 + 1;
  std::vector<std::vector<double>> responses(dirs_azel.size(), std::vector<double>(size_t(nz_max), 0.0));
  for (size_t si = 0; si < dirs_azel.size(); ++si) {
    const mrcsd::Matrix& az_el = dirs_azel[si];
    std::vector<double> cart(size_t(az_el.rows) * 3);
    for (int i = 0; i < az_el.rows; ++i) {   // Math::Sphere::spherical2cartesian (core/math/sphere.h:40-51)
      const double az = az_el(i, 0), el = az_el(i, 1);
      cart[size_t(3 * i)] = std::sin(el) * std::cos(az);
      cart[size_t(3 * i + 1)] = std::sin(el) * std::sin(az);
      cart[size_t(3 * i + 2)] = std::cos(el);
    }
    std::vector<int> vols(volumes[si].begin(), volumes[si].end());
    for (int v : vols)
      if (v >= nvol_img) throw std::runtime_error("direction set has more entries than the image has volumes");
    // the A2SH transform must exist for this shell (amp2response.cpp:305-317): the zonal transform
    // of its elevations is finite (always at lmax 0, where it is the constant P_0; a b=0 shell's
    // zero directions have elevation acos(0 / 0))
    for (int i = 0; i < az_el.rows && lmax[si]; ++i)
      if (!std::isfinite(std::cos(az_el(i, 1)))) {
        std::string msg = "Unable to construct A2SH transformation for shell";
        if (have_shells) msg += " b=" + std::to_string(int(std::lround(shell_list[si].get_mean())));
        msg += ";\n  lmax (" + std::to_string(lmax[si]) + ") may be too large for this shell";
        if (!si && have_shells && shell_list[0].is_bzero())
          msg += "\n  (this appears to be a b=0 shell, and therefore lmax should be set to 0 for this shell)";
        throw std::runtime_error(msg);
      }
    mrresp_ctx* ctx = nullptr;
    check(mrresp_create(cart.data(), vols.data(), int(vols.size()), lmax[si], &ctx));
    const int n = mrresp_ncoef(ctx);
    std::vector<double> M(static_cast<size_t>(n * n)), b(static_cast<size_t>(n)), rf(static_cast<size_t>(n));
    uint64_t count = 0;
    const int rc = mrresp_accumulate(ctx, d_amp, int(nvol_img), d_fib, 3, d_mask, size_t(nspatial), M.data(), b.data(),
                                     &count, nullptr);
    mrresp_destroy(ctx);
    check(rc);
    int iters = 0;
    check(mrresp_solve(lmax[si], M.data(), b.data(), a.noconstraint ? 0 : 1, rf.data(), &iters));
    if (lmax[si] && !a.noconstraint) info("constrained least-squares solver completed in " + std::to_string(iters) + " iterations");
    std::ostringstream line;
    line << "  b=" << (have_shells ? std::to_string(int(std::lround(shell_list[si].get_mean()))) : std::string("?")) << ": [";
    for (int i = 0; i < n; ++i) line << (i ? " " : "") << float(rf[size_t(i)]);
    line << "]";
    console(line.str());
    for (int i = 0; i < n && i < nz_max; ++i) responses[si][size_t(i)] = rf[size_t(i)];
  }
  hip_check(hipFree(d_amp));
  hip_check(hipFree(d_fib));
  hip_check(hipFree(d_mask));
  const auto t2 = std::chrono::steady_clock::now();

  std::ofstream out(a.pos[3]);
  if (!out) throw std::runtime_error("cannot write \"" + a.pos[3] + "\"");
  out.precision(17);
  for (auto& row : responses) {
    for (size_t i = 0; i < row.size(); ++i) out << (i ? " " : "") << row[i];
    out << "\n";
  }
  auto sec = [](auto x, auto y) { return std::chrono::duration<double>(y - x).count(); };
  char buf[160];
  std::snprintf(buf, sizeof buf, "timing: read %.3f s, GPU accumulation and solves incl. transfers %.3f s", sec(t0, t1),
                sec(t1, t2));
  info(buf);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "amp2response: [ERROR] %s\n", e.what());
    return 1;
  }
}
