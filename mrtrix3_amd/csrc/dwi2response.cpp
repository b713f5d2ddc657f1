// dwi2response tournier on the GPU: the script's arguments, checks and iteration
// (bin/dwi2response, lib/mrtrix3/dwi2response/tournier.py:26-154) as one process that keeps the
// DW signal and the FODs on the device between iterations.  Each iteration runs, through
// libmrcsd.so, what the script runs as separate commands:
//
//   dwi2fod csd dwi RF FOD -mask M            mrcsd_run over the voxels of M (include/mrcsd.h)
//   fod2fixel FOD fixel -peak peaks -mask M   mrfixel_segment with no thresholds, its per-voxel
//     -fmls_no_thresholds                     summary (include/mrfixel.h) standing in for
//   fixel2voxel ... none -number 2            the first two fixels' peak amplitudes and
//   fixel2peaks ... -number 1                 the first fixel's direction
//   mrcalc p1 -sqrt 1 p2 p1 -div -sub 2 -pow -mult   the cost, in single precision as mrcalc
//   mrthreshold CF -top N                     the N-th largest non-NaN cost and all at or above
//   amp2response dwi SF dir RF [-lmax]        mrresp_accumulate + mrresp_solve (include/mrresp.h)
//   mrcalc SF iter-1_SF -sub; mrstats max     convergence: no voxel newly selected
//   mrthreshold CF -top 10N | maskfilter dilate -npass D | mrcalc mask - -mult
//                                             the next iteration's mask (6-neighbour dilation)
//
//   dwi2response tournier input output -mask mask [-voxels image] [-shells b] [-lmax l]
//                [-number n] [-iter_voxels n] [-dilate n] [-max_iters n] [-grad file]
//                [-scratch dir] [-nocleanup] [-force] [-quiet] [-info] [-nthreads n]
//
// The input is reduced as `dwiextract -singleshell -no_bzero [-shells b]` does.  -mask is
// required: the brain mask the script computes otherwise (dwi2mask) is not part of this build.
// With -scratch dir -nocleanup, each iteration's RF text file and cost, selection and next-mask
// images are written to dir as the script names them (iter<k>_RF.txt, iter<k>_CF.mif,
// iter<k>_SF.mif, iter<k>_SF_dilated.mif).  Response files are written with 17 significant
// digits and read back exactly.
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
  int number = 300, iter_voxels = 0, dilate = 1, max_iters = 10;
  bool force = false, nocleanup = false;
};

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
      else if (o == "number") a.number = parse_int(o, next());
      else if (o == "iter_voxels") a.iter_voxels = parse_int(o, next());
      else if (o == "dilate") a.dilate = parse_int(o, next());
      else if (o == "max_iters") a.max_iters = parse_int(o, next());
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

int run(int argc, char** argv) {
  const auto t_start = std::chrono::steady_clock::now();
  Args a = parse_args(argc, argv);
  if (a.pos.empty()) throw std::runtime_error("usage: dwi2response algorithm input output [options]");
  if (a.pos[0] != "tournier")
    throw std::runtime_error("algorithm \"" + a.pos[0] + "\" is not part of this build (available: tournier)");
  if (a.pos.size() != 3) throw std::runtime_error("usage: dwi2response tournier input output [options]");
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
  // tournier.py execute(): iteration options
  if (a.max_iters < 2) throw std::runtime_error("Number of iterations must be at least 2");
  int iter_voxels = a.iter_voxels;
  if (iter_voxels == 0) iter_voxels = 10 * a.number;
  else if (iter_voxels < a.number)
    throw std::runtime_error("Number of selected voxels (-iter_voxels) must be greater than number of voxels desired (-number)");

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
  for (int i = 0; i < nvol; ++i)
    for (int j = 0; j < 4; ++j) grad(i, j) = grad_in(int(vols[size_t(i)]), j);
  std::vector<double> grad_flat(size_t(nvol) * 4);
  for (int i = 0; i < nvol; ++i)
    for (int j = 0; j < 4; ++j) grad_flat[size_t(4 * i + j)] = grad(i, j);

  const int64_t* dim = dwi.h.dim.data();
  const int64_t nspatial = dim[0] * dim[1] * dim[2];
  if (a.mask.empty())
    throw std::runtime_error("-mask is required: computing a brain mask (dwi2mask) is not part of this build");
  mrcsd::MifImage mimg = mrcsd::read_image(a.mask);
  if (mimg.h.dim.size() < 3 || mimg.h.dim[0] != dim[0] || mimg.h.dim[1] != dim[1] || mimg.h.dim[2] != dim[2])
    throw std::runtime_error("Dimensions of provided mask image do not match DWI");
  if (!(mimg.h.dim.size() == 3 || (mimg.h.dim.size() == 4 && mimg.h.dim[3] == 1)))
    throw std::runtime_error("Provided mask image needs to be a 3D image");
  std::vector<uint8_t> mask(static_cast<size_t>(nspatial));
  size_t nmask = 0;
  for (int64_t v = 0; v < nspatial; ++v) nmask += (mask[size_t(v)] = mimg.data[size_t(v)] != 0.0f);
  if (!nmask) throw std::runtime_error("Provided mask image does not contain any voxels");
  mrcsd::MifHeader h3 = dwi.h;
  h3.dim.resize(3);
  if (h3.vox.size() > 3) h3.vox.resize(3);
  for (const char* k : {"dw_scheme", "pe_scheme", "PhaseEncodingDirection", "TotalReadoutTime"}) h3.erase(k);

  // the single-shell signal as voxel rows on the device, for every iteration
  std::vector<float> rows(size_t(nspatial) * size_t(nvol));
  for (int64_t v0 = 0; v0 < nspatial; v0 += 256) {
    const int64_t v1 = std::min(nspatial, v0 + 256);
    for (int i = 0; i < nvol; ++i) {
      const float* src = &dwi.data[size_t(int64_t(vols[size_t(i)]) * nspatial)];
      for (int64_t v = v0; v < v1; ++v) rows[size_t(v) * size_t(nvol) + size_t(i)] = src[v];
    }
  }
  hipStream_t st = nullptr;
  hip_check(hipStreamCreate(&st));
  // the grid's rows stay on the device; each iteration packs its mask's voxels (a subset of the
  // initial mask) into the first rows of d_in and runs the CSD, the segmentation and the
  // response accumulation on those only
  const size_t ns = static_cast<size_t>(nspatial);
  Dev<float> d_rows(rows.size()), d_in(nmask * size_t(nvol)), d_summary(nmask * 5);
  Dev<int> d_idx(nmask);
  Dev<uint8_t> d_sf(nmask);
  Dev<uint32_t> d_count(nmask), d_offset(nmask);
  Dev<uint64_t> d_total(1);
  hip_check(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, st));

  // amp2response's shell for dwi.mif: the DW directions of the selected volumes (amp2response.cpp
  // :207-239, spherical2cartesian of gen_direction_matrix)
  std::vector<size_t> all(static_cast<size_t>(nvol));
  for (int i = 0; i < nvol; ++i) all[size_t(i)] = size_t(i);
  const mrcsd::Matrix azel = MR::DWI::gen_direction_matrix(grad, all);
  std::vector<double> cart(size_t(nvol) * 3);
  for (int i = 0; i < nvol; ++i) {
    const double az = azel(i, 0), el = azel(i, 1);
    cart[size_t(3 * i)] = std::sin(el) * std::cos(az);
    cart[size_t(3 * i + 1)] = std::sin(el) * std::sin(az);
    cart[size_t(3 * i + 2)] = std::cos(el);
  }
  std::vector<int> resp_vols(static_cast<size_t>(nvol));
  for (int i = 0; i < nvol; ++i) resp_vols[size_t(i)] = i;

  const bool keep = !a.scratch.empty() && a.nocleanup;
  if (keep && !exists(a.scratch) && ::mkdir(a.scratch.c_str(), 0777) != 0)
    throw std::runtime_error("cannot create scratch directory \"" + a.scratch + "\"");
  auto scratch = [&](const std::string& name) { return a.scratch + "/" + name; };

  std::vector<double> rf = {1.0, -1.0, 1.0};   // init_RF.txt "1 -1 1"
  std::vector<uint8_t> mask_in = mask, sf_prev, sf;
  std::vector<double> final_rf;
  std::vector<float> summary(nmask * 5), cf(ns);
  std::vector<int> idx;
  std::vector<uint8_t> sf_in;
  bool converged = false;
  int iteration = 0;
  // segmenters of the FOD sizes met (lmax 4 in the first iteration, the CSD default after)
  mrfixel_ctx* seg[2] = {nullptr, nullptr};
  int seg_nsh[2] = {-1, -1};
  struct SegFree {
    mrfixel_ctx** s;
    ~SegFree() {
      for (int i = 0; i < 2; ++i)
        if (s[i]) mrfixel_destroy(s[i]);
    }
  } seg_free{seg};
  double fod_ms_total = 0.0;
  for (; iteration < a.max_iters; ++iteration) {
    const auto t_it = std::chrono::steady_clock::now();
    const int amp_lmax = iteration == 0 ? 4 : (lmax.empty() ? -1 : lmax[0]);
    idx.clear();
    for (int64_t v = 0; v < nspatial; ++v)
      if (mask_in[size_t(v)]) idx.push_back(int(v));
    const size_t n_in = idx.size();
    if (!n_in) throw std::runtime_error("no voxel left in the mask to test");
    hip_check(hipMemcpyAsync(d_idx.p, idx.data(), n_in * sizeof(int), hipMemcpyHostToDevice, st));
    {
      const size_t threads = n_in * size_t(nvol);
      gather_rows<<<dim3(unsigned((threads + 255) / 256)), dim3(256), 0, st>>>(d_rows.p, nvol, d_idx.p, n_in, d_in.p);
      hip_check(hipGetLastError());
    }

    // dwi2fod csd dwi.mif RF FOD.mif -mask mask_in
    mrcsd_options copt;
    mrcsd_options_default(&copt);
    mrcsd_shared* shared = nullptr;
    check(mrcsd_shared_create(grad_flat.data(), nvol, rf.data(), int(rf.size()), &copt, &shared));
    mrcsd_shared_info si;
    check(mrcsd_shared_info_get(shared, &si));
    mrcsd_ctx* ctx = nullptr;
    const int rc_ctx = mrcsd_create_from_shared(shared, &ctx);
    mrcsd_shared_destroy(shared);
    check(rc_ctx);
    Dev<float> d_fod(n_in * static_cast<size_t>(si.n_sh));
    hipEvent_t e0, e1, e2;
    hip_check(hipEventCreate(&e0));
    hip_check(hipEventCreate(&e1));
    hip_check(hipEventCreate(&e2));
    hip_check(hipEventRecord(e0, st));
    const int rc_run = mrcsd_run(ctx, d_in.p, nvol, nullptr, n_in, d_fod.p, nullptr, nullptr, st);
    mrcsd_destroy(ctx);
    check(rc_run);
    hip_check(hipEventRecord(e1, st));

    // fod2fixel FOD fixel -peak peaks -mask mask_in -fmls_no_thresholds; fixel2voxel / fixel2peaks
    // (one segmenter per FOD size, kept with its work arrays for the later iterations)
    int slot = seg_nsh[0] == si.n_sh ? 0 : seg_nsh[1] == si.n_sh ? 1 : !seg[0] ? 0 : 1;
    if (seg_nsh[slot] != si.n_sh) {
      if (seg[slot]) mrfixel_destroy(seg[slot]);
      seg[slot] = nullptr;
      seg_nsh[slot] = -1;
      mrfixel_options fopt;
      mrfixel_options_default(&fopt);
      fopt.n_sh = si.n_sh;
      fopt.integral_threshold = 0.0;
      fopt.peak_value_threshold = 0.0;
      check(mrfixel_create(&fopt, &seg[slot]));
      seg_nsh[slot] = si.n_sh;
    }
    mrfixel_ctx* fctx = seg[slot];
    check(mrfixel_segment(fctx, d_fod.p, si.n_sh, nullptr, n_in, d_count.p, d_offset.p, nullptr, 0, d_summary.p,
                          reinterpret_cast<uint64_t*>(d_total.p), st));
    hip_check(hipEventRecord(e2, st));
    hip_check(hipMemcpyAsync(summary.data(), d_summary.p, n_in * 5 * sizeof(float), hipMemcpyDeviceToHost, st));
    hip_check(hipStreamSynchronize(st));
    float ms_csd = 0.0f, ms_fix = 0.0f;
    hip_check(hipEventElapsedTime(&ms_csd, e0, e1));
    hip_check(hipEventElapsedTime(&ms_fix, e1, e2));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    fod_ms_total += ms_csd + ms_fix;

    // the cost sqrt(|p1|) (1 - |p2|/|p1|)^2 in single precision, as mrcalc evaluates it
    // (cmd/mrcalc.cpp:103-123, real_type float); NaN where p1 = 0 (no fixel)
    // (voxels outside this iteration's mask have no fixel: p1 = p2 = 0)
    std::fill(cf.begin(), cf.end(), std::numeric_limits<float>::quiet_NaN());
    size_t voxel_count = 0;
    for (size_t i = 0; i < n_in; ++i) {
      const float p1 = summary[5 * i], p2 = summary[5 * i + 1];
      const float r = 1.0f - p2 / p1;
      const float c = std::sqrt(p1) * std::pow(r, 2.0f);
      cf[size_t(idx[i])] = c;
      voxel_count += std::isfinite(c) ? 1 : 0;   // mrstats count: finite values (core/stats.h:55-70)
    }
    sf = threshold_top(cf, std::min(size_t(a.number), voxel_count));

    // amp2response dwi.mif SF.mif first_dir.mif RF.txt [-lmax]
    const int rl = amp_lmax >= 0 ? amp_lmax : 10;   // amp2response's default for a non-b=0 shell
    if (rl > 12) throw std::runtime_error("lmax must be in [0, 12] on the GPU");
    size_t nsf = 0;
    for (uint8_t s : sf) nsf += s;
    if (!nsf) throw std::runtime_error("input mask does not contain any voxels");
    sf_in.resize(n_in);   // the selection is inside this iteration's mask (the cost is NaN elsewhere)
    for (size_t i = 0; i < n_in; ++i) sf_in[i] = sf[size_t(idx[i])];
    hip_check(hipMemcpyAsync(d_sf.p, sf_in.data(), n_in, hipMemcpyHostToDevice, st));
    mrresp_ctx* rctx = nullptr;
    check(mrresp_create(cart.data(), resp_vols.data(), nvol, rl, &rctx));
    const int n = mrresp_ncoef(rctx);
    std::vector<double> M(static_cast<size_t>(n * n)), b(static_cast<size_t>(n)), rf_new(static_cast<size_t>(n));
    uint64_t count = 0;
    const int rc_acc = mrresp_accumulate(rctx, d_in.p, nvol, d_summary.p + 2, 5, d_sf.p, n_in, M.data(),
                                         b.data(), &count, st);
    mrresp_destroy(rctx);
    check(rc_acc);
    int solver_iters = 0;
    check(mrresp_solve(rl, M.data(), b.data(), 1, rf_new.data(), &solver_iters));
    rf = rf_new;

    {
      std::ostringstream line;
      line << "Optimising (" << iteration + 1 << " iterations, RF: [ ";
      for (size_t i = 0; i < rf.size(); ++i) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%.3f", rf[i]);
        line << (i ? ", " : "") << buf;
      }
      line << "] )";
      console(line.str());
    }
    const std::string prefix = "iter" + std::to_string(iteration) + "_";
    if (keep) {
      write_rf(scratch(prefix + "RF.txt"), rf);
      std::vector<float> cfc = cf;
      mrcsd::write_mif(scratch(prefix + "CF.mif"), h3, cfc, false);
      write_mask(scratch(prefix + "SF.mif"), h3, sf);
    }

    // termination: no voxel newly selected (mrcalc SF prev_SF -sub; mrstats max == 0)
    bool stop = false;
    if (iteration > 0) {
      stop = true;
      for (size_t v = 0; v < sf.size(); ++v)
        if (sf[v] && !sf_prev[v]) {
          stop = false;
          break;
        }
    }
    const double it_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
    char buf[260];
    std::snprintf(buf, sizeof buf,
                  "iteration %d: %zu voxels in the mask; CSD %.2f ms, FMLS %.2f ms (GPU); %zu voxels selected; "
                  "iteration %.3f s, %.3g voxels/s",
                  iteration, n_in, double(ms_csd), double(ms_fix), nsf, it_s, double(n_in) / it_s);
    info(buf);
    if (stop) {
      converged = true;
      break;
    }
    // the next iteration's mask: the top iter_voxels, dilated, within the initial mask
    std::vector<uint8_t> next = dilate(threshold_top(cf, std::min(size_t(iter_voxels), voxel_count)), dim, a.dilate);
    for (size_t v = 0; v < next.size(); ++v) next[v] = next[v] && mask[v];
    if (keep) write_mask(scratch(prefix + "SF_dilated.mif"), h3, next);
    mask_in.swap(next);
    sf_prev = sf;
  }
  if (converged) console("Convergence of SF voxel selection detected at iteration " + std::to_string(iteration + 1));
  else console("Exiting after maximum " + std::to_string(a.max_iters) + " iterations");

  write_rf(output, rf);
  if (!a.voxels.empty()) write_mask(a.voxels, h3, sf);
  hip_check(hipStreamDestroy(st));
  char buf[160];
  std::snprintf(buf, sizeof buf, "timing: %d iterations, GPU CSD + FMLS %.1f ms, total %.3f s",
                converged ? iteration + 1 : a.max_iters, fod_ms_total,
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
