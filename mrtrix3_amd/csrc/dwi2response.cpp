// dwi2response tournier and tax on the GPU: the scripts' arguments, checks and iterations
// (bin/dwi2response, lib/mrtrix3/dwi2response/tournier.py:26-154, tax.py:26-149) as one process
// that keeps the DW signal on the device between iterations.  Each iteration runs, through
// libmrcsd.so, what the scripts run as separate commands:
//
//   dwi2fod csd dwi RF FOD -mask M            mrcsd_run over the voxels of M (include/mrcsd.h)
//   fod2fixel FOD fixel -peak peaks -mask M   mrfixel_segment with no thresholds, its per-voxel
//     -fmls_no_thresholds                     summary (include/mrfixel.h) standing in for
//   fixel2voxel ... none [-number 2]          the first two fixels' peak amplitudes and
//   fixel2peaks ... -number 1                 the first fixel's direction
//   amp2response dwi SF dir RF [-lmax]        mrresp_accumulate + mrresp_solve (include/mrresp.h)
// tournier:
//   mrcalc p1 -sqrt 1 p2 p1 -div -sub 2 -pow -mult   the cost, in single precision as mrcalc
//   mrthreshold CF -top N                     the N-th largest non-NaN cost and all at or above
//   mrcalc SF iter-1_SF -sub; mrstats max     convergence: no voxel newly selected
//   mrthreshold CF -top 10N | maskfilter dilate -npass D | mrcalc mask - -mult
//                                             the next iteration's mask (6-neighbour dilation)
// tax:
//   mrstats dwi -mask -allvolumes             the initial response from the signal's mean and
//                                             standard deviation as mrstats prints them
//   mrcalc p2 p1 -div; ratio R -lt M -mult    the selection, in single precision
//   the response's coefficients all within -convergence percent of the previous ones: converged;
//   else the selection is the next mask
//
//   dwi2response tournier input output -mask mask [-voxels image] [-shells b] [-lmax l]
//                [-number n] [-iter_voxels n] [-dilate n] [-max_iters n] [-grad file]
//                [-scratch dir] [-nocleanup] [-force] [-quiet] [-info] [-nthreads n]
//   dwi2response tax input output -mask mask [-voxels image] [-shells b] [-lmax l]
//                [-peak_ratio r] [-max_iters n] [-convergence c] [-grad file] [...]
//
// The input is reduced as `dwiextract -singleshell -no_bzero [-shells b]` does.  -mask is
// required: the brain mask the script computes otherwise (dwi2mask) is not part of this build.
// With -scratch dir -nocleanup, each iteration's RF text file and selection images are written to
// dir as the scripts name them (iter<k>_RF.txt, iter<k>_SF.mif, and iter<k>_CF.mif,
// iter<k>_SF_dilated.mif for tournier, iter<k>_peak_ratio.mif for tax).  Response files are
// written with 17 significant digits and read back exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mrcsd.h"
#include "../../include/mrfixel.h"
#include "../../include/mrresp.h"
#include "host/gradient.h"
#include "host/mif.h"
#include "host/nifti.h"

namespace {

bool g_quiet = false, g_info = false;

void console(const std::string& s) {
  if (!g_quiet) std::fprintf(stderr, "dwi2response: %s\n", s.c_str());
}
void info(const std::string& s) {
  if (g_info) std::fprintf(stderr, "dwi2response: [INFO] %s\n", s.c_str());
}

void check(int rc) {
  if (rc != MRCSD_OK) throw std::runtime_error(mrcsd_last_error());
}
void hip_check(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
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
  return parse_rows(f);
}

mrcsd::Matrix to_matrix(const std::vector<std::vector<double>>& r, int cols, const std::string& what) {
  mrcsd::Matrix m(int(r.size()), cols);
  for (size_t i = 0; i < r.size(); ++i) {
    if (int(r[i].size()) != cols) throw std::runtime_error(what + ": expected " + std::to_string(cols) + " columns");
    for (int j = 0; j < cols; ++j) m(int(i), j) = r[i][size_t(j)];
  }
  return m;
}

// a positive integer option value (the script's argparse type=int)
int parse_int(const std::string& opt, const std::string& v) {
  size_t used = 0;
  long x = 0;
  try {
    x = std::stol(v, &used);
  } catch (const std::exception&) {
    used = 0;
  }
  if (used != v.size()) throw std::runtime_error("argument -" + opt + ": invalid int value: '" + v + "'");
  return int(x);
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == sep) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  out.push_back(cur);
  return out;
}
struct Args {
  std::vector<std::string> pos;
  std::string mask, voxels, shells, lmax, grad, scratch;
  int number = 300, iter_voxels = 0, dilate = 1, max_iters = -1;   // max_iters: 10 (tournier), 20 (tax)
  double peak_ratio = 0.1, convergence = 0.5;
  std::vector<std::string> given;   // algorithm-specific options given
  bool force = false, nocleanup = false;
};

double parse_float(const std::string& opt, const std::string& v) {
  size_t used = 0;
  double x = 0.0;
  try {
    x = std::stod(v, &used);
  } catch (const std::exception&) {
    used = 0;
  }
  if (used != v.size()) throw std::runtime_error("argument -" + opt + ": invalid float value: '" + v + "'");
  return x;
}

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("argument " + s + ": expected one argument");
      return argv[++i];
    };
    if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1]) && s[1] != '.') {
      const std::string o = s.substr(s[1] == '-' ? 2 : 1);
      if (o == "mask") a.mask = next();
      else if (o == "voxels") a.voxels = next();
      else if (o == "shells") a.shells = next();
      else if (o == "lmax") a.lmax = next();
      else if (o == "grad") a.grad = next();
      else if (o == "fslgrad") throw std::runtime_error("-fslgrad: give the gradient table with -grad or in the image header");
      else if (o == "number") { a.number = parse_int(o, next()); a.given.push_back(o); }
      else if (o == "iter_voxels") { a.iter_voxels = parse_int(o, next()); a.given.push_back(o); }
      else if (o == "dilate") { a.dilate = parse_int(o, next()); a.given.push_back(o); }
      else if (o == "max_iters") a.max_iters = parse_int(o, next());
      else if (o == "peak_ratio") { a.peak_ratio = parse_float(o, next()); a.given.push_back(o); }
      else if (o == "convergence") { a.convergence = parse_float(o, next()); a.given.push_back(o); }
      else if (o == "scratch") a.scratch = next();
      else if (o == "nocleanup") a.nocleanup = true;
      else if (o == "force") a.force = true;
      else if (o == "quiet") g_quiet = true;
      else if (o == "info") g_info = true;
      else if (o == "nthreads") next();
      else throw std::runtime_error("unrecognised option \"" + s + "\"");
    } else {
      a.pos.push_back(s);
    }
  }
  return a;
}


template <class T>
struct Dev {
  T* p = nullptr;
  explicit Dev(size_t n) {
    if (n) hip_check(hipMalloc(&p, n * sizeof(T)));
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
};

// mrthreshold -top n (cmd/mrthreshold.cpp:155-245, 292-299): the (size - n)-th smallest of the
// non-NaN values (nth_element, so the n-th largest) is the threshold, and every non-NaN value at
// or above it is selected
std::vector<uint8_t> threshold_top(const std::vector<float>& cf, size_t n) {
  std::vector<float> data;
  for (float x : cf)
    if (!std::isnan(x)) data.push_back(x);
  if (data.empty()) throw std::runtime_error("No valid input data found; unable to determine threshold");
  const long index = long(data.size()) - long(n);
  if (index < 0 || index >= long(data.size()))
    throw std::runtime_error("Number of valid input image values (" + std::to_string(data.size()) +
                             ") less than number of voxels requested via -top option (" + std::to_string(n) + ")");
  std::nth_element(data.begin(), data.begin() + index, data.end());
  const float thr = data[size_t(index)];
  std::vector<uint8_t> sel(cf.size());
  for (size_t v = 0; v < cf.size(); ++v) sel[v] = !std::isnan(cf[v]) && cf[v] >= thr;
  return sel;
}

// maskfilter dilate -npass n (core/filter/dilate.h:60-104): a voxel is set when it or one of its
// six face neighbours was set in the previous pass; x fastest in m
std::vector<uint8_t> dilate(std::vector<uint8_t> m, const int64_t* dim, int npass) {
  const int64_t nx = dim[0], ny = dim[1], nz = dim[2];
  for (int pass = 0; pass < npass; ++pass) {
    std::vector<uint8_t> out(m.size());
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t v = x + nx * (y + ny * z);
          out[size_t(v)] = m[size_t(v)] || (x > 0 && m[size_t(v - 1)]) || (y > 0 && m[size_t(v - nx)]) ||
                           (z > 0 && m[size_t(v - nx * ny)]) || (x < nx - 1 && m[size_t(v + 1)]) ||
                           (y < ny - 1 && m[size_t(v + nx)]) || (z < nz - 1 && m[size_t(v + nx * ny)]);
        }
    m.swap(out);
  }
  return m;
}

// rows idx[i] of src [*][ld] to dst [n][ld]: this iteration's mask voxels, packed
__global__ void gather_rows(const float* __restrict__ src, int ld, const int* __restrict__ idx, size_t n,
                            float* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t i = t / (size_t)ld;
  const int j = (int)(t - i * (size_t)ld);
  if (i < n) dst[i * (size_t)ld + j] = src[(size_t)idx[i] * (size_t)ld + j];
}

void write_rf(const std::string& path, const std::vector<double>& rf) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write \"" + path + "\"");
  out.precision(17);
  for (size_t i = 0; i < rf.size(); ++i) out << (i ? " " : "") << rf[i];
  out << "\n";
}

void write_mask(const std::string& path, const mrcsd::MifHeader& h3, const std::vector<uint8_t>& m) {
  std::vector<float> d(m.begin(), m.end());
  mrcsd::write_mif(path, h3, d, false);
}

// ostream's default formatting of a double (6 significant digits), read back: how the script sees
// mrstats' mean and standard deviation (core/stats.h:92-98, str<cdouble>, core/mrtrix.h:353-363)
double printed6(double x) {
  char buf[48];
  std::snprintf(buf, sizeof buf, "%g", x);
  return std::strtod(buf, nullptr);
}

// The GPU state of one run: the single-shell rows of the grid on the device, and one iteration's
// analysis of the voxels of a mask -- dwi2fod csd, fod2fixel -fmls_no_thresholds and the first
// two fixels' peak amplitudes and the first fixel's direction -- then amp2response over a
// selection of those voxels.
class Engine {
 public:
  Engine(const std::vector<float>& rows, int nvol, const std::vector<double>& grad_flat, const mrcsd::Matrix& grad,
         size_t nmask)
      : nvol_(nvol), grad_(grad_flat), d_rows_(rows.size()), d_in_(nmask * size_t(nvol)), d_summary_(nmask * 5),
        d_idx_(nmask), d_sf_(nmask), d_count_(nmask), d_offset_(nmask), d_total_(1), summary(nmask * 5) {
    hip_check(hipStreamCreate(&st_));
    hip_check(hipMemcpyAsync(d_rows_.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, st_));
    // amp2response's shell for dwi.mif: the DW directions of the selected volumes (amp2response.cpp
    // :207-239, spherical2cartesian of gen_direction_matrix)
    std::vector<size_t> all(static_cast<size_t>(nvol));
    for (int i = 0; i < nvol; ++i) all[size_t(i)] = size_t(i);
    const mrcsd::Matrix azel = MR::DWI::gen_direction_matrix(grad, all);
    cart_.resize(size_t(nvol) * 3);
    for (int i = 0; i < nvol; ++i) {
      const double az = azel(i, 0), el = azel(i, 1);
      cart_[size_t(3 * i)] = std::sin(el) * std::cos(az);
      cart_[size_t(3 * i + 1)] = std::sin(el) * std::sin(az);
      cart_[size_t(3 * i + 2)] = std::cos(el);
    }
    vols_.resize(static_cast<size_t>(nvol));
    for (int i = 0; i < nvol; ++i) vols_[size_t(i)] = i;
  }
  ~Engine() {
    for (int i = 0; i < 2; ++i)
      if (seg_[i]) mrfixel_destroy(seg_[i]);
    if (st_) (void)hipStreamDestroy(st_);
  }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // the voxels of mask_in (packed in idx order) after CSD with response rf and segmentation:
  // summary [n][5] = first two fixels' peak amplitudes (0 when absent), first fixel's direction
  void analyse(const std::vector<uint8_t>& mask_in, const std::vector<double>& rf) {
    idx.clear();
    for (size_t v = 0; v < mask_in.size(); ++v)
      if (mask_in[v]) idx.push_back(int(v));
    const size_t n = idx.size();
    if (!n) throw std::runtime_error("no voxel left in the mask to test");
    hip_check(hipMemcpyAsync(d_idx_.p, idx.data(), n * sizeof(int), hipMemcpyHostToDevice, st_));
    const size_t threads = n * size_t(nvol_);
    gather_rows<<<dim3(unsigned((threads + 255) / 256)), dim3(256), 0, st_>>>(d_rows_.p, nvol_, d_idx_.p, n, d_in_.p);
    hip_check(hipGetLastError());

    // dwi2fod csd dwi.mif RF FOD.mif -mask mask_in
    mrcsd_options copt;
    mrcsd_options_default(&copt);
    mrcsd_shared* shared = nullptr;
    check(mrcsd_shared_create(grad_.data(), nvol_, rf.data(), int(rf.size()), &copt, &shared));
    mrcsd_shared_info si;
    const int rc_info = mrcsd_shared_info_get(shared, &si);
    mrcsd_ctx* ctx = nullptr;
    const int rc_ctx = rc_info ? rc_info : mrcsd_create_from_shared(shared, &ctx);
    mrcsd_shared_destroy(shared);
    check(rc_ctx);
    Dev<float> d_fod(n * static_cast<size_t>(si.n_sh));
    hipEvent_t e[3];
    for (auto& x : e) hip_check(hipEventCreate(&x));
    hip_check(hipEventRecord(e[0], st_));
    const int rc_run = mrcsd_run(ctx, d_in_.p, nvol_, nullptr, n, d_fod.p, nullptr, nullptr, st_);
    mrcsd_destroy(ctx);
    check(rc_run);
    hip_check(hipEventRecord(e[1], st_));
    // fod2fixel FOD fixel -peak peaks -mask mask_in -fmls_no_thresholds; fixel2voxel / fixel2peaks
    // (one segmenter per FOD size, kept with its work arrays for the later iterations)
    const int slot = seg_nsh_[0] == si.n_sh ? 0 : seg_nsh_[1] == si.n_sh ? 1 : !seg_[0] ? 0 : 1;
    if (seg_nsh_[slot] != si.n_sh) {
      if (seg_[slot]) mrfixel_destroy(seg_[slot]);
      seg_[slot] = nullptr;
      seg_nsh_[slot] = -1;
      mrfixel_options fopt;
      mrfixel_options_default(&fopt);
      fopt.n_sh = si.n_sh;
      fopt.integral_threshold = 0.0;
      fopt.peak_value_threshold = 0.0;
      check(mrfixel_create(&fopt, &seg_[slot]));
      seg_nsh_[slot] = si.n_sh;
    }
    check(mrfixel_segment(seg_[slot], d_fod.p, si.n_sh, nullptr, n, d_count_.p, d_offset_.p, nullptr, 0,
                          d_summary_.p, reinterpret_cast<uint64_t*>(d_total_.p), st_));
    hip_check(hipEventRecord(e[2], st_));
    hip_check(hipMemcpyAsync(summary.data(), d_summary_.p, n * 5 * sizeof(float), hipMemcpyDeviceToHost, st_));
    hip_check(hipStreamSynchronize(st_));
    hip_check(hipEventElapsedTime(&ms_csd, e[0], e[1]));
    hip_check(hipEventElapsedTime(&ms_fmls, e[1], e[2]));
    for (auto& x : e) (void)hipEventDestroy(x);
  }

  // amp2response dwi.mif SF.mif first_dir.mif RF.txt [-lmax l] over the voxels of the last
  // analysis that sf selects (sf over the grid, inside the analysed mask)
  std::vector<double> response(const std::vector<uint8_t>& sf, int lmax) {
    if (lmax < 0 || lmax > 12) throw std::runtime_error("lmax must be in [0, 12] on the GPU");
    const size_t n = idx.size();
    std::vector<uint8_t> sel(n);
    size_t nsel = 0;
    for (size_t i = 0; i < n; ++i) nsel += (sel[i] = sf[size_t(idx[i])]);
    if (!nsel) throw std::runtime_error("input mask does not contain any voxels");
    hip_check(hipMemcpyAsync(d_sf_.p, sel.data(), n, hipMemcpyHostToDevice, st_));
    mrresp_ctx* rctx = nullptr;
    check(mrresp_create(cart_.data(), vols_.data(), nvol_, lmax, &rctx));
    const int nc = mrresp_ncoef(rctx);
    std::vector<double> M(static_cast<size_t>(nc * nc)), b(static_cast<size_t>(nc)), rf(static_cast<size_t>(nc));
    uint64_t count = 0;
    const int rc = mrresp_accumulate(rctx, d_in_.p, nvol_, d_summary_.p + 2, 5, d_sf_.p, n, M.data(), b.data(),
                                     &count, st_);
    mrresp_destroy(rctx);
    check(rc);
    int iters = 0;
    check(mrresp_solve(lmax, M.data(), b.data(), 1, rf.data(), &iters));
    return rf;
  }

  std::vector<int> idx;          // grid indices of the last analysis' voxels
  std::vector<float> summary;    // [idx.size()][5]
  float ms_csd = 0.0f, ms_fmls = 0.0f;

 private:
  int nvol_;
  std::vector<double> grad_, cart_;
  std::vector<int> vols_;
  hipStream_t st_ = nullptr;
  Dev<float> d_rows_, d_in_, d_summary_;
  Dev<int> d_idx_;
  Dev<uint8_t> d_sf_;
  Dev<uint32_t> d_count_, d_offset_;
  Dev<uint64_t> d_total_;
  mrfixel_ctx* seg_[2] = {nullptr, nullptr};
  int seg_nsh_[2] = {-1, -1};
};

std::string rf_text(const std::vector<double>& rf) {
  std::ostringstream line;
  for (size_t i = 0; i < rf.size(); ++i) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.3f", rf[i]);
    line << (i ? ", " : "") << buf;
  }
  return line.str();
}

int run(int argc, char** argv) {
  const auto t_start = std::chrono::steady_clock::now();
  Args a = parse_args(argc, argv);
  if (a.pos.empty()) throw std::runtime_error("usage: dwi2response algorithm input output [options]");
  const std::string algo = a.pos[0];
  if (algo != "tournier" && algo != "tax")
    throw std::runtime_error("algorithm \"" + algo + "\" is not part of this build (available: tax, tournier)");
  for (const auto& o : a.given) {
    const bool tournier_opt = o == "number" || o == "iter_voxels" || o == "dilate";
    if ((algo == "tax") == tournier_opt) throw std::runtime_error("unrecognized arguments: -" + o);
  }
  if (a.pos.size() != 3) throw std::runtime_error("usage: dwi2response " + algo + " input output [options]");
  const std::string& input = a.pos[1];
  const std::string& output = a.pos[2];
  if (exists(output) && !a.force)
    throw std::runtime_error("Output file '" + output + "' already exists (use -force to override)");
  if (!a.voxels.empty() && exists(a.voxels) && !a.force)
    throw std::runtime_error("Output file '" + a.voxels + "' already exists (use -force to override)");

  // bin/dwi2response execute(): -lmax and -shells of a single-shell algorithm
  std::vector<int> lmax;
  if (!a.lmax.empty()) {
    try {   // (an odd value is reported as not a number: the script's bare except)
      for (const auto& t : split(a.lmax, ',')) {
        size_t used = 0;
        const int l = std::stoi(t, &used);
        if (used != t.size() || l % 2) throw std::runtime_error("");
        lmax.push_back(l);
      }
    } catch (const std::exception&) {
      throw std::runtime_error("Parameter lmax must be a number");
    }
    if (lmax.size() != 1) throw std::runtime_error("Can only specify a single lmax value for single-shell algorithms");
  }
  std::vector<double> shells;
  if (!a.shells.empty()) {
    try {
      for (const auto& t : split(a.shells, ',')) {
        size_t used = 0;
        const double b = std::stod(t, &used);
        if (used != t.size()) throw std::runtime_error("");
        shells.push_back(double(std::lround(b)));
      }
    } catch (const std::exception&) {
      throw std::runtime_error("-shells option should provide a comma-separated list of b-values");
    }
    if (shells.size() != 1) throw std::runtime_error("Can only specify a single b-value shell for single-shell algorithms");
  }
  const int max_iters = a.max_iters >= 0 ? a.max_iters : (algo == "tax" ? 20 : 10);
  int iter_voxels = a.iter_voxels;
  if (algo == "tournier") {   // tournier.py execute()
    if (max_iters < 2) throw std::runtime_error("Number of iterations must be at least 2");
    if (iter_voxels == 0) iter_voxels = 10 * a.number;
    else if (iter_voxels < a.number)
      throw std::runtime_error("Number of selected voxels (-iter_voxels) must be greater than number of voxels desired (-number)");
  }

  mrcsd::MifImage dwi = mrcsd::read_image(input);
  if (dwi.h.dim.size() != 4) throw std::runtime_error("input DWI must be a 4-D image");
  const int nvol_in = int(dwi.h.dim[3]);
  std::vector<std::vector<double>> grows;
  if (!a.grad.empty()) {
    grows = load_rows(a.grad);
  } else {
    std::istringstream in(dwi.h.get("dw_scheme"));
    grows = parse_rows(in);
  }
  if (grows.empty())
    throw std::runtime_error("Script requires diffusion gradient table: either in image header, or using -grad / -fslgrad option");
  const mrcsd::Matrix grad_in = MR::DWI::get_DW_scheme(to_matrix(grows, 4, "gradient table"), nvol_in);

  // dwiextract -singleshell -no_bzero [-shells b] (cmd/dwiextract.cpp:70-88)
  MR::DWI::Shells sh(grad_in);
  sh.select_shells(true, false, true, shells);
  std::vector<size_t> vols;
  for (size_t s = 0; s < sh.count(); ++s)
    for (size_t v : sh[s].get_volumes()) vols.push_back(v);
  const int nvol = int(vols.size());
  mrcsd::Matrix grad(nvol, 4);
  std::vector<double> grad_flat(size_t(nvol) * 4);
  for (int i = 0; i < nvol; ++i)
    for (int j = 0; j < 4; ++j) grad_flat[size_t(4 * i + j)] = grad(i, j) = grad_in(int(vols[size_t(i)]), j);

  const int64_t* dim = dwi.h.dim.data();
  const int64_t nspatial = dim[0] * dim[1] * dim[2];
  if (a.mask.empty())
    throw std::runtime_error("-mask is required: computing a brain mask (dwi2mask) is not part of this build");
  mrcsd::MifImage mimg = mrcsd::read_image(a.mask);
  if (mimg.h.dim.size() < 3 || mimg.h.dim[0] != dim[0] || mimg.h.dim[1] != dim[1] || mimg.h.dim[2] != dim[2])
    throw std::runtime_error("Dimensions of provided mask image do not match DWI");
  if (!(mimg.h.dim.size() == 3 || (mimg.h.dim.size() == 4 && mimg.h.dim[3] == 1)))
    throw std::runtime_error("Provided mask image needs to be a 3D image");
  const size_t ns = static_cast<size_t>(nspatial);
  std::vector<uint8_t> mask(ns);
  size_t nmask = 0;
  for (size_t v = 0; v < ns; ++v) nmask += (mask[v] = mimg.data[v] != 0.0f);
  if (!nmask) throw std::runtime_error("Provided mask image does not contain any voxels");
  mrcsd::MifHeader h3 = dwi.h;
  h3.dim.resize(3);
  if (h3.vox.size() > 3) h3.vox.resize(3);
  for (const char* k : {"dw_scheme", "pe_scheme", "PhaseEncodingDirection", "TotalReadoutTime"}) h3.erase(k);

  // the single-shell signal as voxel rows, for every iteration
  std::vector<float> rows(ns * size_t(nvol));
  for (size_t v0 = 0; v0 < ns; v0 += 256) {
    const size_t v1 = std::min(ns, v0 + 256);
    for (int i = 0; i < nvol; ++i) {
      const float* src = &dwi.data[vols[size_t(i)] * ns];
      for (size_t v = v0; v < v1; ++v) rows[v * size_t(nvol) + size_t(i)] = src[v];
    }
  }
  Engine eng(rows, nvol, grad_flat, grad, nmask);

  const bool keep = !a.scratch.empty() && a.nocleanup;
  if (keep && !exists(a.scratch) && ::mkdir(a.scratch.c_str(), 0777) != 0)
    throw std::runtime_error("cannot create scratch directory \"" + a.scratch + "\"");
  auto scratch = [&](const std::string& name) { return a.scratch + "/" + name; };
  auto report = [&](int iteration, const std::chrono::steady_clock::time_point& t_it, size_t nsel) {
    const double it_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
    char buf[260];
    std::snprintf(buf, sizeof buf,
                  "iteration %d: %zu voxels in the mask; CSD %.2f ms, FMLS %.2f ms (GPU); %zu voxels selected; "
                  "iteration %.3f s, %.3g voxels/s",
                  iteration, eng.idx.size(), double(eng.ms_csd), double(eng.ms_fmls), nsel, it_s,
                  double(eng.idx.size()) / it_s);
    info(buf);
  };

  std::vector<double> rf;
  std::vector<uint8_t> mask_in = mask, sf, sf_prev;
  std::vector<float> img(ns);
  bool converged = false;
  int iteration = 0;
  double gpu_ms = 0.0;
  const int amp_lmax = lmax.empty() ? 10 : lmax[0];   // amp2response's default for a non-b=0 shell

  if (algo == "tournier") {
    rf = {1.0, -1.0, 1.0};   // init_RF.txt "1 -1 1"
    for (; iteration < max_iters; ++iteration) {
      const auto t_it = std::chrono::steady_clock::now();
      eng.analyse(mask_in, rf);
      gpu_ms += eng.ms_csd + eng.ms_fmls;
      // the cost sqrt(p1) (1 - p2/p1)^2 in single precision, as mrcalc evaluates it
      // (cmd/mrcalc.cpp:103-123, real_type float); NaN without a fixel, and outside the mask
      std::fill(img.begin(), img.end(), std::numeric_limits<float>::quiet_NaN());
      size_t voxel_count = 0;
      for (size_t i = 0; i < eng.idx.size(); ++i) {
        const float p1 = eng.summary[5 * i], p2 = eng.summary[5 * i + 1];
        const float r = 1.0f - p2 / p1;
        const float c = std::sqrt(p1) * std::pow(r, 2.0f);
        img[size_t(eng.idx[i])] = c;
        voxel_count += std::isfinite(c) ? 1 : 0;   // mrstats count: finite values (core/stats.h:55-70)
      }
      sf = threshold_top(img, std::min(size_t(a.number), voxel_count));
      rf = eng.response(sf, iteration == 0 ? 4 : amp_lmax);
      console("Optimising (" + std::to_string(iteration + 1) + " iterations, RF: [ " + rf_text(rf) + "] )");
      const std::string prefix = "iter" + std::to_string(iteration) + "_";
      if (keep) {
        write_rf(scratch(prefix + "RF.txt"), rf);
        mrcsd::write_mif(scratch(prefix + "CF.mif"), h3, img, false);
        write_mask(scratch(prefix + "SF.mif"), h3, sf);
      }
      // termination: no voxel newly selected (mrcalc SF prev_SF -sub; mrstats max == 0)
      bool stop = iteration > 0;
      for (size_t v = 0; stop && v < ns; ++v)
        if (sf[v] && !sf_prev[v]) stop = false;
      size_t nsf = 0;
      for (uint8_t s : sf) nsf += s;
      report(iteration, t_it, nsf);
      if (stop) {
        converged = true;
        break;
      }
      // the next iteration's mask: the top iter_voxels, dilated, within the initial mask
      std::vector<uint8_t> next = dilate(threshold_top(img, std::min(size_t(iter_voxels), voxel_count)), dim, a.dilate);
      for (size_t v = 0; v < ns; ++v) next[v] = next[v] && mask[v];
      if (keep) write_mask(scratch(prefix + "SF_dilated.mif"), h3, next);
      mask_in.swap(next);
      sf_prev = sf;
    }
    if (converged) console("Convergence of SF voxel selection detected at iteration " + std::to_string(iteration + 1));
    else console("Exiting after maximum " + std::to_string(max_iters) + " iterations");
  } else {
    // tax.py execute(): the initial response from the mean and standard deviation of the DW signal
    // in the mask (mrstats -allvolumes -mask: finite values, Welford's running mean and variance)
    double mean = 0.0, m2 = 0.0;
    size_t count = 0;
    for (size_t v = 0; v < ns; ++v) {
      if (!mask[v]) continue;
      for (int i = 0; i < nvol; ++i) {
        const double x = rows[v * size_t(nvol) + size_t(i)];
        if (!std::isfinite(x)) continue;
        ++count;
        const double delta = x - mean;
        mean += delta / double(count);
        m2 += delta * (x - mean);
      }
    }
    const double std_dev = count > 1 ? std::sqrt(m2 / double(count - 1)) : 0.0;
    const double s4pi = std::sqrt(4.0 * M_PI);
    const double m = printed6(mean) * s4pi, s = printed6(std_dev) * s4pi;
    rf = {m, -0.5 * s, 0.25 * s * s / m};
    if (keep) write_rf(scratch("init_RF.txt"), rf);
    const double change = 0.01 * a.convergence;
    const float thr = float(a.peak_ratio);
    size_t nsf = 0;
    for (; iteration < max_iters; ++iteration) {
      const auto t_it = std::chrono::steady_clock::now();
      eng.analyse(mask_in, rf);
      gpu_ms += eng.ms_csd + eng.ms_fmls;
      // peak_ratio = second / first in single precision; SF = (ratio < peak_ratio) within the mask
      std::fill(img.begin(), img.end(), std::numeric_limits<float>::quiet_NaN());
      sf.assign(ns, 0);
      nsf = 0;
      for (size_t i = 0; i < eng.idx.size(); ++i) {
        const float r = eng.summary[5 * i + 1] / eng.summary[5 * i];
        img[size_t(eng.idx[i])] = r;
        nsf += (sf[size_t(eng.idx[i])] = r < thr);
      }
      if (!nsf) throw std::runtime_error("Aborting: All voxels have been excluded from single-fibre selection");
      const std::vector<double> rf_new = eng.response(sf, amp_lmax);
      console("Optimising (" + std::to_string(iteration + 1) + " iterations, " + std::to_string(nsf) +
              " voxels, RF: [ " + rf_text(rf_new) + "] )");
      const std::string prefix = "iter" + std::to_string(iteration) + "_";
      if (keep) {
        write_rf(scratch(prefix + "RF.txt"), rf_new);
        mrcsd::write_mif(scratch(prefix + "peak_ratio.mif"), h3, img, false);
        write_mask(scratch(prefix + "SF.mif"), h3, sf);
      }
      // termination: every coefficient's half-difference within `change` of its mean
      bool reiterate = iteration == 0;
      for (size_t k = 0; !reiterate && k < std::min(rf.size(), rf_new.size()); ++k) {
        const double mid = 0.5 * (rf[k] + rf_new[k]);
        const double diff = std::fabs(0.5 * (rf[k] - rf_new[k]));
        if (diff / mid > change) reiterate = true;
      }
      rf = rf_new;
      report(iteration, t_it, nsf);
      if (!reiterate) {
        converged = true;
        break;
      }
      mask_in = sf;
    }
    if (converged)
      console("Exited at iteration " + std::to_string(iteration + 1) + " with " + std::to_string(nsf) +
              " SF voxels due to unchanged RF coefficients");
    else
      console("Exited after maximum " + std::to_string(max_iters) + " iterations with " + std::to_string(nsf) +
              " SF voxels");
  }

  write_rf(output, rf);
  if (!a.voxels.empty()) write_mask(a.voxels, h3, sf);
  char buf[160];
  std::snprintf(buf, sizeof buf, "timing: %d iterations, GPU CSD + FMLS %.1f ms, total %.3f s",
                converged ? iteration + 1 : max_iters, gpu_ms,
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count());
  info(buf);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "dwi2response: [ERROR] %s\n", e.what());
    return 1;
  }
}
