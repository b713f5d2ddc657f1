// Runs mricls on a case written by tests/test_icls_kernel_emulated.py through the C ABI, with the
// ICLS kernel source compiled for the CPU against emu_runtime.h (KERNEL_SOURCE: the kernel file
// without its launcher) and run one workgroup at a time.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/mricls.h"
#include KERNEL_SOURCE

namespace mrcsd {
hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, hipStream_t) {
  *L.counter = 0ull;
  const unsigned grid = 2;
  for (unsigned b = 0; b < grid; ++b) emu::run_workgroup([&]() { icls_persistent_kernel(P, L); }, b, grid);
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

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string d = argv[1];
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
