// Runs mricls on a case written by tests/test_icls_kernel_emulated.py through the C ABI, with the
// ICLS kernel source compiled for the CPU against emu_runtime.h (KERNEL_SOURCE: the kernel file
// without its launcher) and run one workgroup at a time.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/mricls.h"
#include KERNEL_SOURCE

namespace mrcsd {
// The passes of icls_launch, each as two emulated persistent workgroups run one after the other.
template <int KCAP, int KQ>
void run_pass(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsPass& Q) {
  const unsigned grid = 2;
  for (unsigned b = 0; b < grid; ++b)
    emu::run_workgroup([&]() { icls_active_set_kernel<KCAP, KQ, false>(P, L, Q); }, b, grid);
}
template <int KCAP>
void run_kq(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsPass& Q) {
  switch (P.ncon_pad / 64) {
    case 4: run_pass<KCAP, 4>(P, L, Q); break;
    case 5: run_pass<KCAP, 5>(P, L, Q); break;
    case 8: run_pass<KCAP, 8>(P, L, Q); break;
    default: run_pass<KCAP, 16>(P, L, Q); break;
  }
}
hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsScratch& S, int first,
                       hipStream_t) {
  int caps[kIclsMaxPasses];
  const int np = plan_passes(first, caps);
  for (int i = 0; i < 2 * kIclsMaxPasses; ++i) S.counters[i] = 0ull;
  for (int i = 0; i < np; ++i) {
    const IclsPass Q = pass_args(S, i, np);
    switch (caps[i]) {
      case 16: run_kq<16>(P, L, Q); break;
      case 32: run_kq<32>(P, L, Q); break;
      case 48: run_kq<48>(P, L, Q); break;
      case 64: run_kq<64>(P, L, Q); break;
      case 96: run_kq<96>(P, L, Q); break;
      case 128: run_kq<128>(P, L, Q); break;
      default: return hipErrorInvalidValue;
    }
    std::fprintf(stderr, "pass %d: capacity %d, queued for the next pass %llu\n", i, caps[i],
                 i + 1 < np ? (unsigned long long)S.counters[kIclsMaxPasses + i] : 0ull);
  }
  return hipSuccess;
}
// mricls_predict's launcher, restated on the host (the emulation runs the active-set kernel only)
hipError_t icls_predict(const double* Ht, int n, int nmeas, const float* x, int ld_x, size_t nvox, float* out,
                        int ld_out, hipStream_t) {
  for (size_t v = 0; v < nvox; ++v)
    for (int r = 0; r < nmeas; ++r) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc = std::fma(Ht[size_t(j) * nmeas + r], (double)x[v * size_t(ld_x) + j], acc);
      out[v * size_t(ld_out) + r] = (float)acc;
    }
  return hipSuccess;
}
}  // namespace mrcsd

template <class T>
std::vector<T> load(const std::string& p, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(p, std::ios::binary);
  f.read((char*)v.data(), std::streamsize(n * sizeof(T)));
  if (!f) {
    std::fprintf(stderr, "cannot read %s\n", p.c_str());
    std::exit(2);
  }
  return v;
}

// A plain ICLS problem (generic.txt: nmeas n ncon nvox con_reg; H.bin, A.bin, b.bin).
int run_generic(const std::string& d, int capacity) {
  int nm, n, nc, nvox;
  double con_reg;
  std::ifstream h(d + "/generic.txt");
  h >> nm >> n >> nc >> nvox >> con_reg;
  auto H = load<double>(d + "/H.bin", size_t(nm) * size_t(n));
  auto A = load<double>(d + "/A.bin", size_t(nc) * size_t(n));
  auto b = load<float>(d + "/b.bin", size_t(nvox) * size_t(nm));
  mricls_ctx* ctx = nullptr;
  if (mricls_create(H.data(), nm, n, A.data(), nc, nullptr, 0, 0.0, con_reg, 0, 0.0, 0, &ctx) ||
      (capacity && mricls_set_capacity(ctx, capacity))) {
    std::fprintf(stderr, "icls: %s\n", mrcsd_last_error());
    return 2;
  }
  std::vector<float> x(size_t(nvox) * size_t(n));
  std::vector<unsigned short> it(static_cast<size_t>(nvox)), na(static_cast<size_t>(nvox));
  if (mricls_solve_host(ctx, b.data(), nm, nullptr, size_t(nvox), x.data(), it.data(), na.data())) {
    std::fprintf(stderr, "solve: %s\n", mrcsd_last_error());
    return 2;
  }
  std::ofstream out(d + "/x.bin", std::ios::binary);
  out.write((const char*)x.data(), std::streamsize(x.size() * sizeof(float)));
  std::ofstream o2(d + "/iters.txt");
  for (int v = 0; v < nvox; ++v) o2 << it[size_t(v)] << " " << na[size_t(v)] << "\n";
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string d = argv[1];
  if (std::ifstream(d + "/generic.txt").good()) return run_generic(d, argc > 2 ? std::atoi(argv[2]) : 0);
  int nvol, nvox, ld, nt;
  std::ifstream h(d + "/case.txt");
  h >> nvol >> nvox >> ld >> nt;
  std::vector<int> rows(static_cast<size_t>(nt)), cols(static_cast<size_t>(nt));
  size_t nresp = 0;
  for (int i = 0; i < nt; ++i) {
    h >> rows[size_t(i)] >> cols[size_t(i)];
    nresp += size_t(rows[size_t(i)]) * size_t(cols[size_t(i)]);
  }
  auto grad = load<double>(d + "/grad.bin", size_t(nvol) * 4);
  auto resp = load<double>(d + "/resp.bin", nresp);
  auto dw = load<float>(d + "/dw.bin", size_t(nvox) * size_t(ld));
  auto valid = load<unsigned char>(d + "/valid.bin", size_t(nvox));
  mrmsmt_shared* sh = nullptr;
  if (mrmsmt_shared_create(grad.data(), nvol, resp.data(), rows.data(), cols.data(), nt, nullptr, 0, nullptr, 0, -1.0,
                           -1.0, nullptr, 0, &sh)) {
    std::fprintf(stderr, "shared: %s\n", mrcsd_last_error());
    return 2;
  }
  mricls_ctx* ctx = nullptr;
  if (mrmsmt_icls_create(sh, &ctx)) {
    std::fprintf(stderr, "icls: %s\n", mrcsd_last_error());
    return 2;
  }
  if (argc > 2 && mricls_set_capacity(ctx, std::atoi(argv[2]))) {
    std::fprintf(stderr, "capacity: %s\n", mrcsd_last_error());
    return 2;
  }
  int nm, n, nc, mi;
  mricls_dims(ctx, &nm, &n, &nc, &mi);
  std::vector<float> x(size_t(nvox) * size_t(n));
  std::vector<unsigned short> it(static_cast<size_t>(nvox)), na(static_cast<size_t>(nvox));
  if (mricls_solve_host(ctx, dw.data(), ld, valid.data(), size_t(nvox), x.data(), it.data(), na.data())) {
    std::fprintf(stderr, "solve: %s\n", mrcsd_last_error());
    return 2;
  }
  std::ofstream out(d + "/x.bin", std::ios::binary);
  out.write((const char*)x.data(), std::streamsize(x.size() * sizeof(float)));
  std::ofstream o2(d + "/iters.txt");
  for (int v = 0; v < nvox; ++v) o2 << it[size_t(v)] << " " << na[size_t(v)] << "\n";
  std::printf("n %d collectives %lld\n", n, emu::g_collectives);
  return 0;
}
