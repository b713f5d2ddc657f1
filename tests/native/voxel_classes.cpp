// Test driver of the per-voxel classes (host/csd.h, host/msmt.h, host/icls.h) in the shape of the
// reference's processors: dwi2fod's CSD_Processor (cmd/dwi2fod.cpp:120-176) and MSMT_Processor
// (:181-231) run over the voxels by several threads, each with its own copy of the solver object
// (ThreadedLoop copies the functor per thread, core/thread.h:148-158), handing out voxels under a
// lock as ThreadedLoop hands out lines.  Built by mrtrix3_amd/build.py; run by
// tests/test_gpu_voxel_classes.py (GPU) and tests/test_voxel_classes_cpu.py (no GPU).
//
//   voxel_classes csd  <in> <out> <threads>   in: int32 nvol, nvox, n_resp, lmax; double grad[nvol][4],
//                                                 response[n_resp]; float dw[nvox][nvol]
//                                             out: int32 n_sh; float fod[nvox][n_sh]; int32 n[nvox]
//   voxel_classes trajectory <in> <out> <cmax> for voxel v < nvox and c = 0..cmax: set, c calls of
//                                             iterate() (stopping early when one returns true), FOD()
//                                             out: int32 n_sh; float fod[nvox][cmax + 1][n_sh]
//   voxel_classes msmt <in> <out> <threads>   in: int32 nvol, nvox, ntissue; int32 rows[t], cols[t],
//                                                 lmax[t]; double grad[nvol][4], responses (row-major,
//                                                 concatenated); float dw[nvox][nvol]
//                                             out: int32 n; double x[nvox][n]; int32 niter[nvox]
//   voxel_classes nogpu                       constructs the classes without a GPU: prints the
//                                             messages of the exceptions they throw
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "host/csd.h"
#include "host/gradient.h"
#include "host/icls.h"
#include "host/msmt.h"

using MR::DWI::SDeconv::CSD;
using MR::DWI::SDeconv::MSMT_CSD;
using mrcsd::Matrix;

namespace {

struct Reader {
  std::ifstream f;
  explicit Reader(const char* path) : f(path, std::ios::binary) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  }
  template <class T>
  std::vector<T> get(size_t n) {
    std::vector<T> v(n);
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T)));
    if (!f) throw std::runtime_error("truncated input");
    return v;
  }
  int32_t i32() { return get<int32_t>(1)[0]; }
};

template <class T>
void put(std::ofstream& o, const std::vector<T>& v) {
  o.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

// Voxels handed out one at a time under a lock (ThreadedLoop's outer-loop handout).
struct Handout {
  std::mutex m;
  size_t next = 0, n = 0;
  bool take(size_t& v) {
    std::lock_guard<std::mutex> l(m);
    if (next >= n) return false;
    v = next++;
    return true;
  }
};

struct CsdInput {
  int nvol = 0, nvox = 0;
  std::unique_ptr<CSD::Shared> shared;
  std::vector<float> dw;
};

CsdInput read_csd(const char* path) {
  Reader r(path);
  CsdInput in;
  in.nvol = r.i32();
  in.nvox = r.i32();
  const int n_resp = r.i32(), lmax = r.i32();
  auto g = r.get<double>(size_t(in.nvol) * 4);
  auto resp = r.get<double>(size_t(n_resp));
  in.dw = r.get<float>(size_t(in.nvox) * in.nvol);
  Matrix raw(in.nvol, 4);
  raw.a = g;
  in.shared = std::make_unique<CSD::Shared>(MR::DWI::get_DW_scheme(raw, in.nvol));
  MR::DWI::SDeconv::CSDOptions opt;
  if (lmax) opt.lmax.push_back(uint32_t(lmax));
  in.shared->parse_cmdline_options(opt);
  in.shared->set_response(resp);
  in.shared->init();
  return in;
}

// CSD_Processor::load_data (dwi2fod.cpp:156-173): the selected volumes, false on a non-finite
// value, negatives set to 0.
bool load_data(const CSD::Shared& s, const float* row, std::vector<double>& out) {
  out.resize(s.dwis.size());
  for (size_t n = 0; n < s.dwis.size(); ++n) {
    const float v = row[s.dwis[n]];
    if (!std::isfinite(v)) return false;
    out[n] = v < 0.0f ? 0.0 : double(v);
  }
  return true;
}

int run_csd(const char* in_path, const char* out_path, int threads) {
  CsdInput in = read_csd(in_path);
  const CSD::Shared& S = *in.shared;
  const size_t nsh = S.nSH();
  std::vector<float> fod(size_t(in.nvox) * nsh, 0.0f);
  std::vector<int32_t> iters(size_t(in.nvox), -1);
  CSD proto(S);
  Handout h;
  h.n = size_t(in.nvox);
  std::vector<std::thread> pool;
  std::vector<std::string> errors(static_cast<size_t>(threads));
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      try {
        CSD csd(proto);   // the per-thread copy
        std::vector<double> dwi;
        size_t v;
        while (h.take(v)) {
          if (!load_data(S, &in.dw[v * size_t(in.nvol)], dwi)) continue;   // zeros, as dwi2fod writes
          csd.set(dwi);
          size_t n;
          for (n = 0; n < S.niter; n++)
            if (csd.iterate()) break;
          const auto& F = csd.FOD();
          for (size_t i = 0; i < nsh; ++i) fod[v * nsh + i] = float(F[i]);
          iters[v] = int32_t(n);
        }
      } catch (const std::exception& e) {
        errors[size_t(t)] = e.what();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errors)
    if (!e.empty()) {
      std::fprintf(stderr, "error: %s\n", e.c_str());
      return 1;
    }
  std::ofstream o(out_path, std::ios::binary);
  put(o, std::vector<int32_t>{int32_t(nsh)});
  put(o, fod);
  put(o, iters);
  return 0;
}

int run_trajectory(const char* in_path, const char* out_path, int cmax) {
  CsdInput in = read_csd(in_path);
  const CSD::Shared& S = *in.shared;
  const size_t nsh = S.nSH();
  std::vector<float> out(size_t(in.nvox) * size_t(cmax + 1) * nsh, 0.0f);
  CSD csd(S);
  std::vector<double> dwi;
  for (size_t v = 0; v < size_t(in.nvox); ++v) {
    if (!load_data(S, &in.dw[v * size_t(in.nvol)], dwi)) continue;
    for (int c = 0; c <= cmax; ++c) {
      csd.set(dwi);
      for (int k = 0; k < c; ++k)
        if (csd.iterate()) break;
      const auto& F = csd.FOD();
      for (size_t i = 0; i < nsh; ++i) out[(v * size_t(cmax + 1) + size_t(c)) * nsh + i] = float(F[i]);
    }
  }
  std::ofstream o(out_path, std::ios::binary);
  put(o, std::vector<int32_t>{int32_t(nsh)});
  put(o, out);
  return 0;
}

int run_msmt(const char* in_path, const char* out_path, int threads) {
  Reader r(in_path);
  const int nvol = r.i32(), nvox = r.i32(), nt = r.i32();
  auto rows = r.get<int32_t>(size_t(nt)), cols = r.get<int32_t>(size_t(nt)), lmax = r.get<int32_t>(size_t(nt));
  auto g = r.get<double>(size_t(nvol) * 4);
  std::vector<Matrix> resp;
  for (int t = 0; t < nt; ++t) {
    Matrix m(rows[size_t(t)], cols[size_t(t)]);
    m.a = r.get<double>(size_t(rows[size_t(t)]) * size_t(cols[size_t(t)]));
    resp.push_back(m);
  }
  auto dw = r.get<float>(size_t(nvox) * nvol);
  Matrix raw(nvol, 4);
  raw.a = g;
  MSMT_CSD::Shared S(MR::DWI::get_DW_scheme(raw, nvol));
  S.lmax.assign(lmax.begin(), lmax.end());
  S.set_responses(resp);
  S.init();
  const size_t n = size_t(S.C.cols);
  std::vector<double> x(size_t(nvox) * n, 0.0);
  std::vector<int32_t> niter(size_t(nvox), -1);
  MSMT_CSD proto(S);
  Handout h;
  h.n = size_t(nvox);
  std::vector<std::thread> pool;
  std::vector<std::string> errors(static_cast<size_t>(threads));
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      try {
        MSMT_CSD sdeconv(proto);
        std::vector<double> data(static_cast<size_t>(nvol)), output;
        size_t v;
        while (h.take(v)) {
          // MSMT_Processor (dwi2fod.cpp:196-214): every volume of the voxel, as read
          for (int k = 0; k < nvol; ++k) data[size_t(k)] = double(dw[v * size_t(nvol) + size_t(k)]);
          sdeconv(data, output);
          for (size_t i = 0; i < n; ++i) x[v * n + i] = output[i];
          niter[v] = int32_t(sdeconv.niter);
        }
      } catch (const std::exception& e) {
        errors[size_t(t)] = e.what();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errors)
    if (!e.empty()) {
      std::fprintf(stderr, "error: %s\n", e.c_str());
      return 1;
    }
  std::ofstream o(out_path, std::ios::binary);
  put(o, std::vector<int32_t>{int32_t(n)});
  put(o, x);
  put(o, niter);
  return 0;
}

int run_nogpu() {
  // a single-shell scheme: 6 b=0 and 60 b=3000 volumes on a spiral
  const int nvol = 66;
  Matrix raw(nvol, 4);
  for (int i = 0; i < nvol; ++i) {
    if (i < 6) continue;
    const double z = 1.0 - 2.0 * (i - 6 + 0.5) / 60.0, rr = std::sqrt(1.0 - z * z), ph = 2.399963 * i;
    raw(i, 0) = rr * std::cos(ph);
    raw(i, 1) = rr * std::sin(ph);
    raw(i, 2) = z;
    raw(i, 3) = 3000.0;
  }
  CSD::Shared S(MR::DWI::get_DW_scheme(raw, nvol));
  S.parse_cmdline_options(MR::DWI::SDeconv::CSDOptions());
  S.set_response(std::vector<double>{1000.0, -400.0, 100.0, -20.0, 3.0});
  S.init();
  try {
    CSD csd(S);
    std::printf("CSD constructed\n");
  } catch (const std::exception& e) {
    std::printf("CSD: %s\n", e.what());
  }
  Matrix H(3, 2), A(2, 2);
  H.a = {1, 0, 0, 1, 1, 1};
  A.a = {1, 0, 0, 1};
  MR::Math::ICLS::Problem P(H, A);
  try {
    MR::Math::ICLS::Solver solver(P);
    std::printf("Solver constructed\n");
  } catch (const std::exception& e) {
    std::printf("Solver: %s\n", e.what());
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "nogpu") return run_nogpu();
    if (argc != 5) {
      std::fprintf(stderr, "usage: voxel_classes csd|trajectory|msmt <in> <out> <n> | nogpu\n");
      return 2;
    }
    const int k = std::atoi(argv[4]);
    if (mode == "csd") return run_csd(argv[2], argv[3], k);
    if (mode == "trajectory") return run_trajectory(argv[2], argv[3], k);
    if (mode == "msmt") return run_msmt(argv[2], argv[3], k);
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
