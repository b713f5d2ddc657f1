// dwi2fod — command-line drop-in for MRtrix3's dwi2fod (cmd/dwi2fod.cpp) running the voxel fits
// on the GPU through the C ABI of libmrcsd.so (include/mrcsd.h, include/mricls.h).
//
//   dwi2fod csd dwi.mif response.txt fod.mif [options]
//   dwi2fod msmt_csd dwi.mif wm.txt wm.mif gm.txt gm.mif csf.txt csf.mif [options]
//
// Options kept from the reference: -mask, -lmax, -directions, -filter, -neg_lambda,
// -norm_lambda, -threshold, -niter, -grad, -fslgrad, -shells, -bvalue_scaling, -predicted_signal, -force, -quiet,
// -info, -nthreads (accepted; the solve runs on the GPU).  Per-voxel semantics follow
// dwi2fod.cpp:128-231: masked-out voxels are written as zeros, CSD voxels with a non-finite
// signal give zeros, negative CSD signals are clamped, MSMT uses the raw signal.
// Added: -warm_net net.npz (csd): start each voxel from the negative set proposed by a network
// trained for the acquisition protocol (include/mrwarm.h); a voxel is accepted only when that set
// is a fixed point of the iteration, else it is solved from the reference's start.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mrcsd.h"
#include "../../include/mricls.h"
#include "../../include/mrwarm.h"
#include "host/mif.h"
#include "host/nifti.h"

namespace {

bool g_quiet = false, g_info = false;

void console(const std::string& s) {
  if (!g_quiet) std::fprintf(stderr, "dwi2fod: %s\n", s.c_str());
}
void info(const std::string& s) {
  if (g_info) std::fprintf(stderr, "dwi2fod: [INFO] %s\n", s.c_str());
}

// Wall time of each stage (image read, solver set-up, GPU solve including the host/device
// transfers, image write), reported with -info as one "timing:" line.
struct Stages {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now(), last = t0;
  std::string text;
  void mark(const char* name) {
    const auto now = std::chrono::steady_clock::now();
    char buf[96];
    std::snprintf(buf, sizeof buf, "%s%s %.3f s", text.empty() ? "" : ", ", name,
                  std::chrono::duration<double>(now - last).count());
    text += buf;
    last = now;
  }
  void report() {
    char buf[64];
    std::snprintf(buf, sizeof buf, ", total %.3f s",
                  std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    info("timing: " + text + buf);
  }
};

std::vector<std::vector<double>> parse_rows(const std::string& text) {
  std::vector<std::vector<double>> rows;
  std::istringstream in(text);
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
  std::stringstream ss;
  ss << f.rdbuf();
  auto r = parse_rows(ss.str());
  if (r.empty()) throw std::runtime_error("no data in file \"" + path + "\"");
  for (auto& row : r)
    if (row.size() != r[0].size()) throw std::runtime_error("uneven rows in file \"" + path + "\"");
  return r;
}

std::vector<double> flatten(const std::vector<std::vector<double>>& r) {
  std::vector<double> out;
  for (auto& row : r) out.insert(out.end(), row.begin(), row.end());
  return out;
}

std::vector<int> parse_ints(const std::string& s) {
  std::vector<int> out;
  std::string cur;
  for (char c : s + ",") {
    if (c == ',') {
      if (!cur.empty()) out.push_back(std::stoi(cur));
      cur.clear();
    } else {
      cur += c;
    }
  }
  return out;
}

std::string scheme_text(const std::vector<double>& g, int nvol) {
  std::ostringstream o;
  o.precision(17);
  for (int r = 0; r < nvol; ++r) {
    if (r) o << "\n";
    o << g[size_t(r) * 4] << "," << g[size_t(r) * 4 + 1] << "," << g[size_t(r) * 4 + 2] << "," << g[size_t(r) * 4 + 3];
  }
  return o.str();
}

void check(int rc) {
  if (rc != MRCSD_OK) throw std::runtime_error(mrcsd_last_error());
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device memory of the command's solve (freed on scope exit).
template <class T>
struct Dev {
  T* p = nullptr;
  explicit Dev(size_t n) {
    if (n) HIPCHK(hipMalloc(&p, n * sizeof(T)));
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { HIPCHK(hipStreamCreate(&s)); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  ~Stream() { (void)hipStreamDestroy(s); }
};

struct Args {
  std::vector<std::string> pos;
  std::string mask, directions, filter, grad, predicted, warm_net;
  std::vector<std::string> fslgrad;   // bvecs, bvals
  std::vector<double> shells;         // -shells b-values
  std::vector<int> lmax;
  bool has_neg = false, has_norm = false, has_thr = false, has_niter = false, force = false;
  double neg = 0, norm = 0, thr = 0;
  int niter = 50, bscale = 0;
};

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("no value supplied for option \"" + s + "\"");
      return argv[++i];
    };
    if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1])) {
      std::string o = s.substr(s[1] == '-' ? 2 : 1);
      if (o == "mask") a.mask = next();
      else if (o == "lmax") a.lmax = parse_ints(next());
      else if (o == "directions") a.directions = next();
      else if (o == "filter") a.filter = next();
      else if (o == "grad") a.grad = next();
      else if (o == "fslgrad") {
        const std::string bvecs = next();
        a.fslgrad = {bvecs, next()};
      }
      else if (o == "predicted_signal") a.predicted = next();
      else if (o == "warm_net") a.warm_net = next();
      else if (o == "shells") {
        std::string cur;
        for (char c : next() + ",") {
          if (c == ',') {
            if (!cur.empty()) a.shells.push_back(std::stod(cur));
            cur.clear();
          } else if (c != ' ') {
            cur += c;
          }
        }
      }
      else if (o == "neg_lambda") { a.has_neg = true; a.neg = std::stod(next()); }
      else if (o == "norm_lambda") { a.has_norm = true; a.norm = std::stod(next()); }
      else if (o == "threshold") { a.has_thr = true; a.thr = std::stod(next()); }
      else if (o == "niter") { a.has_niter = true; a.niter = std::stoi(next()); }
      else if (o == "bvalue_scaling") {
        const std::string v = next();
        a.bscale = (v == "yes" || v == "true" || v == "1") ? 1 : 2;
      }
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

bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// Gathers the voxels to solve: returns their canonical (x-fastest) indices and packs their
// signals as rows [nsel][nvol] (the DWI buffer is in canonical order with axis 0 fastest).
void gather(const mrcsd::MifImage& dwi, const std::vector<float>* mask, std::vector<int64_t>& sel,
            std::vector<float>& rows) {
  const int64_t nspatial = dwi.h.dim[0] * dwi.h.dim[1] * dwi.h.dim[2];
  const int64_t nvol = dwi.h.dim[3];
  sel.clear();
  for (int64_t v = 0; v < nspatial; ++v)
    if (!mask || (*mask)[size_t(v)] != 0.0f) sel.push_back(v);
  rows.resize(sel.size() * size_t(nvol));
  // blocks of 256 voxels: the reads of one volume and the block's rows both stay in cache
  for (size_t i0 = 0; i0 < sel.size(); i0 += 256) {
    const size_t i1 = std::min(sel.size(), i0 + 256);
    for (int64_t c = 0; c < nvol; ++c) {
      const float* src = &dwi.data[size_t(c * nspatial)];
      for (size_t i = i0; i < i1; ++i) rows[i * size_t(nvol) + size_t(c)] = src[sel[i]];
    }
  }
}

mrcsd::MifHeader output_header(const mrcsd::MifHeader& in, int64_t nvol_out, const std::vector<double>& grad,
                               int nvol) {
  mrcsd::MifHeader h = in;
  h.dim = {in.dim[0], in.dim[1], in.dim[2], nvol_out};
  h.vox.resize(3);
  h.vox.push_back(1.0);
  h.erase("dw_scheme");
  h.set("prior_dw_scheme", scheme_text(grad, nvol));
  for (const char* k : {"pe_scheme", "PhaseEncodingDirection", "TotalReadoutTime"}) h.erase(k);
  return h;
}

void scatter_write(const std::string& path, const mrcsd::MifHeader& h, const std::vector<int64_t>& sel,
                   const float* rows, int ncoef, int ld, int off) {
  const int64_t nspatial = h.dim[0] * h.dim[1] * h.dim[2];
  std::vector<float> out(size_t(nspatial * ncoef), 0.0f);
  for (size_t i0 = 0; i0 < sel.size(); i0 += 256) {
    const size_t i1 = std::min(sel.size(), i0 + 256);
    for (int c = 0; c < ncoef; ++c) {
      float* dst = &out[size_t(c) * size_t(nspatial)];
      for (size_t i = i0; i < i1; ++i) dst[sel[i]] = rows[i * size_t(ld) + size_t(off + c)];
    }
  }
  mrcsd::write_mif(path, h, out, true);
}

// core/dwi/gradient.cpp:127-200 (load_bvecs_bvals): an FSL bvecs / bvals pair as a gradient
// table in scanner space.  bvecs are given in the image axes of a left-handed system: the first
// component is negated when the stored transform's 3x3 part has a positive determinant, then the
// vectors are rotated by it; rows with NaN become b = 0.
std::vector<double> load_bvecs_bvals(const mrcsd::MifHeader& h, int64_t nvol, const std::string& bvecs_path,
                                     const std::string& bvals_path) {
  std::vector<std::vector<double>> bvals, bvecs;
  try {
    bvals = load_rows(bvals_path);
    bvecs = load_rows(bvecs_path);
  } catch (const std::exception& e) {
    throw std::runtime_error(std::string(e.what()) + "\nUnable to import files \"" + bvecs_path + "\" and \"" +
                             bvals_path + "\" as FSL bvecs/bvals pair");
  }
  auto transpose = [](const std::vector<std::vector<double>>& m) {
    std::vector<std::vector<double>> t(m[0].size(), std::vector<double>(m.size()));
    for (size_t r = 0; r < m.size(); ++r)
      for (size_t c = 0; c < m[r].size(); ++c) t[c][r] = m[r][c];
    return t;
  };
  if (bvals.size() != 1) {
    if (bvals[0].size() == 1) bvals = transpose(bvals);
    else
      throw std::runtime_error("bvals file must contain 1 row or column only (file \"" + bvals_path + "\" has " +
                               std::to_string(bvals.size()) + ")");
  }
  if (bvecs.size() != 3) {
    if (bvecs[0].size() == 3) bvecs = transpose(bvecs);
    else
      throw std::runtime_error("bvecs file must contain exactly 3 rows or columns (file \"" + bvecs_path + "\" has " +
                               std::to_string(bvecs.size()) + ")");
  }
  const size_t n = bvals[0].size();
  if (bvecs[0].size() != n)
    throw std::runtime_error("bvecs and bvals files must have same number of diffusion directions (file \"" + bvecs_path +
                             "\" has " + std::to_string(bvecs[0].size()) + ", file \"" + bvals_path + "\" has " +
                             std::to_string(n) + ")");
  if (int64_t(n) != nvol)
    throw std::runtime_error("bvecs and bvals files must have same number of diffusion directions as DW-image (gradients: " +
                             std::to_string(n) + ", image: " + std::to_string(nvol) + ")");
  double L[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (h.transform.size() >= 3)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) L[r][c] = h.transform[size_t(r)][size_t(c)];
  const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) - L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                     L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
  if (det > 0.0)
    for (auto& v : bvecs[0]) v = -v;
  std::vector<double> grad(n * 4);
  size_t nan_rows = 0;
  bool nan_bvecs = false, nan_bvals = false;
  for (size_t i = 0; i < n; ++i) {
    double g[3];
    for (int r = 0; r < 3; ++r) g[r] = L[r][0] * bvecs[0][i] + L[r][1] * bvecs[1][i] + L[r][2] * bvecs[2][i];
    double b = bvals[0][i];
    const bool nan_b = std::isnan(b), nan_v = std::isnan(g[0]) || std::isnan(g[1]) || std::isnan(g[2]);
    if (nan_b && !nan_v && g[0] * g[0] + g[1] * g[1] + g[2] * g[2] > 0.0)
      throw std::runtime_error("Corrupt content in bvecs/bvals data (" + bvecs_path + " & " + bvals_path +
                               ") (NaN present in bval but valid direction in bvec)");
    if (nan_v && b > 0.0)
      throw std::runtime_error("Corrupt content in bvecs/bvals data (" + bvecs_path + " & " + bvals_path +
                               ") (NaN bvec direction but non-zero value in bval)");
    nan_bvals |= nan_b;
    nan_bvecs |= nan_v;
    if (nan_b || nan_v) {
      g[0] = g[1] = g[2] = 0.0;
      b = 0.0;
      ++nan_rows;
    }
    grad[i * 4] = g[0];
    grad[i * 4 + 1] = g[1];
    grad[i * 4 + 2] = g[2];
    grad[i * 4 + 3] = b;
  }
  if (nan_rows)
    console("[WARNING] " + std::to_string(nan_rows) + " row" + (nan_rows > 1 ? "s" : "") + " with NaN values detected in " +
            (nan_bvecs ? "bvecs file " + bvecs_path + (nan_bvals ? " and" : "") : std::string()) +
            (nan_bvals ? "bvals file " + bvals_path : std::string()) + "; these have been interpreted as b=0 volumes by MRtrix");
  return grad;
}

int run(int argc, char** argv) {
  Args a = parse_args(argc, argv);
  if (a.pos.size() < 2) throw std::runtime_error("usage: dwi2fod algorithm dwi response odf [response odf ...]");
  const std::string algo = a.pos[0];
  if (algo != "csd" && algo != "msmt_csd")
    throw std::runtime_error("algorithm must be one of: csd,msmt_csd");
  Stages stages;
  console("reading image \"" + a.pos[1] + "\"");
  mrcsd::MifImage dwi = mrcsd::read_image(a.pos[1]);
  if (dwi.h.dim.size() != 4) throw std::runtime_error("input DWI must be a 4-D image");
  const int nvol = int(dwi.h.dim[3]);

  std::vector<double> grad;
  if (!a.grad.empty() && !a.fslgrad.empty())   // gradient.cpp:261-265
    throw std::runtime_error("Diffusion gradient table can be provided using either -grad or -fslgrad option, but NOT both");
  if (!a.grad.empty()) grad = flatten(load_rows(a.grad));
  else if (!a.fslgrad.empty()) grad = load_bvecs_bvals(dwi.h, nvol, a.fslgrad[0], a.fslgrad[1]);
  else grad = flatten(parse_rows(dwi.h.get("dw_scheme")));
  if (grad.empty()) throw std::runtime_error("no valid diffusion gradient table found");
  if (grad.size() != size_t(nvol) * 4)
    throw std::runtime_error("number of studies in base image (" + std::to_string(nvol) +
                             ") does not match number of rows in diffusion gradient table (" +
                             std::to_string(grad.size() / 4) + ")");

  std::vector<float> maskv;
  if (!a.mask.empty()) {
    mrcsd::MifImage m = mrcsd::read_image(a.mask);
    if (m.h.dim.size() < 3 || m.h.dim[0] != dwi.h.dim[0] || m.h.dim[1] != dwi.h.dim[1] || m.h.dim[2] != dwi.h.dim[2])
      throw std::runtime_error("dimension mismatch between \"" + a.pos[1] + "\" and mask \"" + a.mask + "\"");
    maskv.assign(m.data.begin(), m.data.begin() + dwi.h.dim[0] * dwi.h.dim[1] * dwi.h.dim[2]);
  }
  std::vector<int64_t> sel;
  std::vector<float> rows;
  gather(dwi, a.mask.empty() ? nullptr : &maskv, sel, rows);
  const size_t nsel = sel.size();
  stages.mark("read");

  if (algo == "csd") {
    if (a.pos.size() != 4)
      throw std::runtime_error("CSD algorithm expects a single input response function and single output FOD image");
    if (exists(a.pos[3]) && !a.force) throw std::runtime_error("output image \"" + a.pos[3] + "\" already exists (use -force option to force overwrite)");
    std::vector<double> response;
    try {
      auto r = load_rows(a.pos[2]);
      if (r.size() != 1 && r[0].size() != 1) throw std::runtime_error("response file contains matrix, not vector");
      response = flatten(r);
    } catch (const std::exception& e) {
      throw std::runtime_error(std::string(e.what()) + "\nCSD algorithm expects second argument to be the input response function file");
    }
    mrcsd_options opt;
    mrcsd_options_default(&opt);
    if (a.lmax.size() > 1) throw std::runtime_error("CSD algorithm expects a single lmax to be specified");
    if (!a.lmax.empty()) opt.lmax = a.lmax[0];
    std::vector<double> filt, dirs;
    if (!a.filter.empty()) { filt = flatten(load_rows(a.filter)); opt.filter = filt.data(); opt.n_filter = int(filt.size()); }
    if (!a.directions.empty()) {
      auto d = load_rows(a.directions);
      if (d[0].size() != 2) throw std::runtime_error("direction matrix should have 2 columns: [ azimuth elevation ]");
      dirs = flatten(d);
      opt.directions = dirs.data();
      opt.n_directions = int(d.size());
    }
    if (a.has_neg) opt.neg_lambda = a.neg;
    if (a.has_norm) opt.norm_lambda = a.norm;
    if (a.has_thr) opt.threshold = a.thr;
    if (a.has_niter) opt.niter = a.niter;
    opt.bvalue_scaling = a.bscale;
    if (!a.shells.empty()) { opt.shells = a.shells.data(); opt.n_shells = int(a.shells.size()); }
    mrcsd_shared* shared = nullptr;
    check(mrcsd_shared_create(grad.data(), nvol, response.data(), int(response.size()), &opt, &shared));
    mrcsd_shared_info si;
    check(mrcsd_shared_info_get(shared, &si));
    std::vector<double> gused(size_t(nvol) * 4);
    check(mrcsd_shared_matrix(shared, 5, gused.data(), gused.size()));
    mrcsd_ctx* ctx = nullptr;
    check(mrcsd_create_from_shared(shared, &ctx));
    stages.mark("setup");
    console("performing constrained spherical deconvolution on " + std::to_string(nsel) + " voxels (GPU)");
    std::vector<float> fod(nsel * size_t(si.n_sh));
    std::vector<uint16_t> iters(nsel);
    bool warm = !a.warm_net.empty();
    if (warm && !mrcsd_warm_supported(ctx)) {
      console("[WARNING] -warm_net ignored: the warm start's fixed-point test certifies the reference's answer "
              "only for -threshold 0 and a positive definite normal matrix (-norm_lambda > 0)");
      warm = false;
    }
    if (!warm) {
      check(mrcsd_run_host(ctx, rows.data(), nvol, nullptr, nsel, fod.data(), iters.data(), nullptr));
    } else {
      mrwarm_net* net = nullptr;
      check(mrwarm_load(a.warm_net.c_str(), &net));
      std::unique_ptr<mrwarm_net, void (*)(mrwarm_net*)> hold(net, mrwarm_destroy);
      int n_in = 0, hidden = 0, nsh = 0, nlog = 0;
      check(mrwarm_dims(net, &n_in, &hidden, &nsh, &nlog));
      if (n_in != si.n_dw || nsh != si.n_sh || nlog != si.n_hr)
        throw std::runtime_error("network \"" + a.warm_net + "\" was trained for " + std::to_string(n_in) +
                                 " DW volumes, " + std::to_string(nsh) + " coefficients and " + std::to_string(nlog) +
                                 " constraint directions; this run has " + std::to_string(si.n_dw) + ", " +
                                 std::to_string(si.n_sh) + " and " + std::to_string(si.n_hr));
      std::vector<double> dwd(size_t(si.n_dw));
      check(mrcsd_shared_matrix(shared, 4, dwd.data(), dwd.size()));
      std::vector<int> dwis(dwd.begin(), dwd.end()), b0;
      for (int r = 0; r < nvol; ++r)
        if (gused[size_t(r) * 4 + 3] <= 10.0) b0.push_back(r);   // bzero_threshold() (gradient.cpp:92-95)
      const size_t words = size_t(si.n_hr + 63) / 64;
      Stream st;
      Dev<float> d_rows(nsel * size_t(nvol)), d_fod(nsel * size_t(si.n_sh));
      Dev<int> d_dwis(dwis.size()), d_b0(b0.size());
      Dev<uint64_t> d_neg(nsel * words), d_stats(6);
      Dev<uint16_t> d_it(nsel);
      HIPCHK(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, st.s));
      HIPCHK(hipMemcpyAsync(d_dwis.p, dwis.data(), dwis.size() * sizeof(int), hipMemcpyHostToDevice, st.s));
      if (!b0.empty())
        HIPCHK(hipMemcpyAsync(d_b0.p, b0.data(), b0.size() * sizeof(int), hipMemcpyHostToDevice, st.s));
      const uint64_t zero[6] = {0, 0, 0, 0, 0, 0};
      HIPCHK(hipMemcpyAsync(d_stats.p, zero, sizeof zero, hipMemcpyHostToDevice, st.s));
      check(mrwarm_infer(net, d_rows.p, nvol, nsel, d_dwis.p, d_b0.p, int(b0.size()), nullptr, d_neg.p, st.s));
      mrcsd_outputs o{d_it.p, nullptr, nullptr, d_stats.p};
      mrcsd_warm w{nullptr, d_neg.p, 6};
      check(mrcsd_run(ctx, d_rows.p, nvol, nullptr, nsel, d_fod.p, &o, &w, st.s));
      uint64_t stats[6];
      HIPCHK(hipMemcpyAsync(fod.data(), d_fod.p, fod.size() * sizeof(float), hipMemcpyDeviceToHost, st.s));
      HIPCHK(hipMemcpyAsync(iters.data(), d_it.p, nsel * sizeof(uint16_t), hipMemcpyDeviceToHost, st.s));
      HIPCHK(hipMemcpyAsync(stats, d_stats.p, sizeof stats, hipMemcpyDeviceToHost, st.s));
      HIPCHK(hipStreamSynchronize(st.s));
      info("warm start: " + std::to_string(stats[5]) + " solves from the proposals, " + std::to_string(stats[4]) +
           " voxels restarted from the plain initialisation, " + std::to_string(stats[1]) + " solves in all (" +
           std::to_string(stats[0]) + " voxels)");
    }
    stages.mark("GPU solve incl. transfers");
    size_t nonconv = 0;
    for (size_t i = 0; i < nsel; ++i)
      if (iters[i] != 0xFFFF && si.niter && iters[i] >= si.niter) ++nonconv;
    if (nonconv) info(std::to_string(nonconv) + " voxels did not reach full convergence");
    mrcsd::MifHeader h = output_header(dwi.h, si.n_sh, gused, nvol);
    scatter_write(a.pos[3], h, sel, fod.data(), si.n_sh, si.n_sh, 0);
    stages.mark("write");
    stages.report();
    mrcsd_destroy(ctx);
    mrcsd_shared_destroy(shared);
    return 0;
  }

  // msmt_csd
  if (a.pos.size() % 2)
    throw std::runtime_error("MSMT_CSD algorithm expects pairs of (input response function & output FOD image) to be provided");
  const int nt = int(a.pos.size() - 2) / 2;
  std::vector<double> resp;
  std::vector<int> rrows, rcols;
  for (int t = 0; t < nt; ++t) {
    const std::string& f = a.pos[size_t(2 + 2 * t)];
    std::vector<std::vector<double>> r;
    try {
      r = load_rows(f);
    } catch (const std::exception& e) {
      throw std::runtime_error("File \"" + f + "\" is not a valid response function file\n"
                               "MSMT_CSD algorithm expects the first file in each argument pair to be an input response function file");
    }
    rrows.push_back(int(r.size()));
    rcols.push_back(int(r[0].size()));
    auto fl = flatten(r);
    resp.insert(resp.end(), fl.begin(), fl.end());
    if (exists(a.pos[size_t(3 + 2 * t)]) && !a.force)
      throw std::runtime_error("output image \"" + a.pos[size_t(3 + 2 * t)] + "\" already exists (use -force option to force overwrite)");
  }
  std::vector<double> dirs;
  int ndirs = 0;
  if (!a.directions.empty()) {
    auto d = load_rows(a.directions);
    dirs = flatten(d);
    ndirs = int(d.size());
  }
  mrmsmt_shared* ms = nullptr;
  check(mrmsmt_shared_create(grad.data(), nvol, resp.data(), rrows.data(), rcols.data(), nt,
                             a.lmax.empty() ? nullptr : a.lmax.data(), int(a.lmax.size()),
                             dirs.empty() ? nullptr : dirs.data(), ndirs,
                             a.has_norm ? a.norm : -1.0, a.has_neg ? a.neg : -1.0,
                             a.shells.empty() ? nullptr : a.shells.data(), int(a.shells.size()), &ms));
  int nshell = 0, nparams = 0, ncon = 0;
  std::vector<int> lm(static_cast<size_t>(nt));
  check(mrmsmt_shared_info(ms, &nshell, &nparams, &ncon, lm.data()));
  mricls_ctx* ic = nullptr;
  check(mrmsmt_icls_create(ms, &ic));
  stages.mark("setup");
  console("performing MSMT CSD (" + std::to_string(nshell) + " shell" + (nshell > 1 ? "s" : "") + ", " +
          std::to_string(nt) + " tissue" + (nt > 1 ? "s" : "") + ") on " + std::to_string(nsel) + " voxels (GPU)");
  std::vector<float> x(nsel * size_t(nparams)), pred;
  std::vector<uint16_t> iters(nsel);
  {
    // signals up, the ICLS solve and the predicted signal H x (dwi2fod.cpp:216-220) on the GPU,
    // results down
    Stream st;
    Dev<float> d_rows(nsel * size_t(nvol)), d_x(nsel * size_t(nparams));
    Dev<uint16_t> d_it(nsel);
    HIPCHK(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice, st.s));
    check(mricls_solve(ic, d_rows.p, nvol, nullptr, nsel, d_x.p, d_it.p, nullptr, nullptr, st.s));
    HIPCHK(hipMemcpyAsync(x.data(), d_x.p, x.size() * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIPCHK(hipMemcpyAsync(iters.data(), d_it.p, nsel * sizeof(uint16_t), hipMemcpyDeviceToHost, st.s));
    if (!a.predicted.empty()) {
      check(mricls_predict(ic, d_x.p, nparams, nsel, d_rows.p, nvol, st.s));   // over the signals
      pred.resize(rows.size());
      HIPCHK(hipMemcpyAsync(pred.data(), d_rows.p, pred.size() * sizeof(float), hipMemcpyDeviceToHost, st.s));
    }
    HIPCHK(hipStreamSynchronize(st.s));
  }
  const size_t over = size_t(std::count(iters.begin(), iters.end(), uint16_t(0xFFFE)));
  if (over)
    throw std::runtime_error(std::to_string(over) + " voxel(s) need more than 128 active constraints, beyond the GPU "
                             "solver's capacity");
  stages.mark("GPU solve incl. transfers");
  std::vector<double> gused = grad;   // stash the table as interpreted (normalised) by the set-up
  {
    for (int r = 0; r < nvol; ++r) {
      double* g = &gused[size_t(r) * 4];
      const double sq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
      if (sq > 0) { const double n = std::sqrt(sq); g[0] /= n; g[1] /= n; g[2] /= n; }
    }
  }
  int off = 0;
  for (int t = 0; t < nt; ++t) {
    const int ncoef = (lm[size_t(t)] + 1) * (lm[size_t(t)] + 2) / 2;
    mrcsd::MifHeader h = output_header(dwi.h, ncoef, gused, nvol);
    scatter_write(a.pos[size_t(3 + 2 * t)], h, sel, x.data(), ncoef, nparams, off);
    off += ncoef;
  }
  if (!a.predicted.empty()) {
    mrcsd::MifHeader h = dwi.h;
    scatter_write(a.predicted, h, sel, pred.data(), nvol, nvol, 0);
  }
  stages.mark("write");
  stages.report();
  mricls_destroy(ic);
  mrmsmt_shared_destroy(ms);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "dwi2fod: [ERROR] %s\n", e.what());
    return 1;
  }
}
