// Runs mrfixel on a case written by tests/test_fmls_kernel_emulated.py through the C ABI, with the
// FMLS sort, growth, revision, offset and emit kernels compiled for the CPU against emu_runtime.h
// (KERNEL_SOURCE: the kernel file without its launchers) and run one workgroup at a time.  The
// amplitude GEMM (FMA in ascending coefficient order, as the GPU kernel accumulates, into the same
// voxel-major layout) is restated here on the host.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/mrfixel.h"
// the growth kernels' dynamic LDS: a buffer of exactly the launch's size (so that AddressSanitizer
// checks the columns' bounds), shared by the lanes of the running workgroup
static unsigned char* emu_grow_lds = nullptr;
#define FMLS_GROW_LDS_ARRAY emu_grow_lds
#include KERNEL_SOURCE

namespace mrcsd {
template <class F>
void run_grid(unsigned grid, F&& body) {
  for (unsigned b = 0; b < grid; ++b) emu::run_workgroup([&]() { body(); }, b, grid);
}

// the revision over a grid of three workgroups, so that lanes stride over several items
template <int LMAX>
void run_revise(const FmlsDeviceSet& S, const FmlsLaunch& L, size_t v0) {
  run_grid(3, [&]() { fmls_revise_kernel<LMAX>(S, L, v0); });
}

// the launch sequence of fmls_launch, each kernel run one workgroup at a time
hipError_t fmls_launch(const FmlsDeviceSet& S, const FmlsLaunch& L, hipStream_t) {
  if (S.ndir < 64 || S.ndir > kFmlsMaxDirs || L.chunk == 0 || L.chunk % kFmlsBlock) return hipErrorInvalidValue;
  *L.total = 0ull;
  for (size_t v0 = 0; v0 < L.nvox; v0 += L.chunk) {
    const size_t nv = L.nvox - v0 < L.chunk ? L.nvox - v0 : L.chunk;
    const unsigned nb = unsigned((nv + kFmlsBlock - 1) / kFmlsBlock);
    const size_t lda = size_t(fmls_amp_ld(S.ndir));
    for (size_t v = 0; v < nv; ++v)
      for (int d = 0; d < S.ndir; ++d) {
        double acc = 0.0;
        for (int k = 0; k < S.nsh; ++k)
          acc = std::fma(S.sht[size_t(d) * S.nsh + k], (double)L.fod[(v0 + v) * size_t(L.ld_fod) + k], acc);
        L.work.amps[v * lda + d] = acc;
      }
    run_grid(unsigned(nv), [&]() { fmls_sort_kernel(S, L, v0, nv); });
    std::vector<unsigned char> lds8(size_t(S.ndir) * S.lanes8), lds16(size_t(S.ndir) * S.lanes16 * 2);
    emu_grow_lds = lds8.data();
    run_grid(unsigned((nv + S.lanes8 - 1) / S.lanes8), [&]() { fmls_grow_kernel<Lab8>(S, L, nv, S.lanes8, 0); });
    emu_grow_lds = lds16.data();
    run_grid(unsigned((nv + S.lanes16 - 1) / S.lanes16), [&]() { fmls_grow_kernel<Lab16>(S, L, nv, S.lanes16, 1); });
    run_grid(nb, [&]() { fmls_block_sums(reinterpret_cast<const uint32_t*>(L.work.nlobes), 0, nv, L.sums); });
    *L.work.nitems = 0ull;
    run_grid(1, [&]() { fmls_scan_sums(L.sums, nb, L.work.nitems); });
    run_grid(nb, [&]() { fmls_lobe_items(L, nv); });
    switch (S.lmax) {
      case 2: run_revise<2>(S, L, v0); break;
      case 4: run_revise<4>(S, L, v0); break;
      case 6: run_revise<6>(S, L, v0); break;
      case 8: run_revise<8>(S, L, v0); break;
      case 10: run_revise<10>(S, L, v0); break;
      case 12: run_revise<12>(S, L, v0); break;
      default: return hipErrorInvalidValue;
    }
    run_grid(nb, [&]() { fmls_finalize_kernel(S, L, v0, nv); });
    run_grid(nb, [&]() { fmls_block_sums(L.count, v0, nv, L.sums); });
    run_grid(1, [&]() { fmls_scan_sums(L.sums, nb, L.total); });
    run_grid(nb, [&]() { fmls_emit(L, v0, nv); });
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

template <class T>
void save(const std::string& p, const T* v, size_t n) {
  std::ofstream f(p, std::ios::binary);
  f.write((const char*)v, std::streamsize(n * sizeof(T)));
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string d = argv[1];
  // case.txt: nvox n_sh maxnum dir_from_peak / integral_threshold peak_threshold merge_ratio
  size_t nvox;
  int n_sh, maxnum, dirpeak;
  double ithr, pthr, ratio;
  std::ifstream h(d + "/case.txt");
  h >> nvox >> n_sh >> maxnum >> dirpeak >> ithr >> pthr >> ratio;
  if (!h) return 2;
  auto fod = load<float>(d + "/fod.bin", nvox * size_t(n_sh));
  auto mask = load<unsigned char>(d + "/mask.bin", nvox);
  mrfixel_options opt;
  mrfixel_options_default(&opt);
  opt.n_sh = n_sh;
  opt.maxnum = maxnum;
  opt.dir_from_peak = dirpeak;
  opt.integral_threshold = ithr;
  opt.peak_value_threshold = pthr;
  opt.lobe_merge_ratio = ratio;
  mrfixel_ctx* ctx = nullptr;
  if (mrfixel_create(&opt, &ctx)) {
    std::fprintf(stderr, "create: %s\n", mrcsd_last_error());
    return 2;
  }
  const size_t cap = 64 * nvox;
  std::vector<uint32_t> count(nvox), offset(nvox);
  std::vector<float> fixels(5 * cap), summary(5 * nvox);
  uint64_t total = 0;
  if (mrfixel_segment_host(ctx, fod.data(), n_sh, mask.data(), nvox, count.data(), offset.data(), fixels.data(), cap,
                           summary.data(), &total)) {
    std::fprintf(stderr, "segment: %s\n", mrcsd_last_error());
    return 2;
  }
  save(d + "/count.bin", count.data(), nvox);
  save(d + "/offset.bin", offset.data(), nvox);
  save(d + "/fixels.bin", fixels.data(), 5 * size_t(total));
  save(d + "/summary.bin", summary.data(), 5 * nvox);
  std::printf("total %llu collectives %lld\n", (unsigned long long)total, emu::g_collectives);
  mrfixel_destroy(ctx);
  return 0;
}
