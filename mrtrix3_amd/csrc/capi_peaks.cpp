// C ABI of the GPU peak finder (include/mrpeaks.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrpeaks.h"
#include "capi_common.h"
#include "host/sh.h"
#include "kernels/peaks.h"

namespace {

const float kDefaultSeeds[] = {
#include "host/seeds60.inc"
};

// sh2peaks.cpp:93-97 (Direction(phi, theta)): the unit vector, in single precision
void unit_vector(float phi, float theta, float* v) {
  v[0] = std::cos(phi) * std::sin(theta);
  v[1] = std::sin(phi) * std::sin(theta);
  v[2] = std::cos(theta);
}

}  // namespace

struct mrpeaks_ctx {
  int device = 0;
  int n_sh = 0, lmax = 0, npeaks = 0, nseeds = 0;
  float threshold = -INFINITY;
  bool fixed_dirs = false;
  mrcsd::DeviceBuffer<float> seeds;     // [nseeds][3]
  mrcsd::DeviceBuffer<float> dirs;      // [npeaks][3] (-direction)
  mrcsd::DeviceBuffer<unsigned long long> counter;   // the search kernel's work counter
  mrcsd::DeviceBuffer<float4> work;                  // search results, grown on demand
  mrcsd::DeviceBuffer<float> fast_table;             // -fast lookup table, or empty
};

// Legendre::Plm_sph from cos(elevation) alone (legendre.h:98-122), in single precision with the
// double steps of the reference: P_l^m for l = m..lmax into a[l].
void plm_sph_x(float* a, int lmax, int m, float x) {
  const float x2 = x * x;
  if (m && double(x2) >= 1.0) {
    for (int n = m; n <= lmax; ++n) a[n] = 0.0f;
    return;
  }
  a[m] = float(0.282094791773878);
  if (m) {
    double h = 1.0;   // Plm_sph_helper(1 - x^2, 2m) (legendre.h:66-69)
    for (double k = 2.0; k <= 2.0 * m; k += 2.0) h = (1.0 - double(x2)) * (k - 1.0) / k * h;   // innermost first
    a[m] = float(double(a[m]) * std::sqrt(double(float(2 * m + 1)) * h));
  }
  if (m & 1) a[m] = -a[m];
  if (lmax == m) return;
  float f = std::sqrt(float(2 * m + 3));
  a[m + 1] = x * f * a[m];
  for (int n = m + 2; n <= lmax; ++n) {
    a[n] = x * a[n - 1] - a[n - 2] / f;
    f = std::sqrt(float(4 * n * n - 1) / float(n * n - m * m));
    a[n] *= f;
  }
}

// PrecomputedAL::init (SH.h:418-434) with sh2peaks' 512 elevations: row n holds the values at
// elevation n pi / 511, entry index_mpos(l, m) = l^2/4 + m.
std::vector<float> fast_table(int lmax) {
  const int nal = (lmax / 2 + 1) * (lmax / 2 + 1);
  const float inc = float(M_PI / (mrcsd::kPeaksFastRows - 1));
  std::vector<float> t(size_t(mrcsd::kPeaksFastRows) * size_t(nal)), buf(size_t(lmax) + 1);
  for (int n = 0; n < mrcsd::kPeaksFastRows; ++n) {
    const float cos_el = std::cos(float(n) * inc);
    for (int m = 0; m <= lmax; ++m) {
      plm_sph_x(buf.data(), lmax, m, cos_el);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) t[size_t(n) * nal + size_t(l * l / 4 + m)] = buf[size_t(l)];
    }
  }
  return t;
}

extern "C" {

void mrpeaks_options_default(mrpeaks_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->npeaks = 3;                 // DEFAULT_NPEAKS (sh2peaks.cpp:26)
  opt->threshold = -INFINITY;      // sh2peaks.cpp:335
}

int mrpeaks_create(const mrpeaks_options* opt, mrpeaks_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!opt || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_create: null argument");
    *out = nullptr;
    const int lmax = mrcsd::sh::LforN(opt->n_sh);
    if (opt->n_sh < 1 || opt->n_sh > 128 || lmax < 2 || lmax > 12)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_create: the SH series must have lmax 2 to 12 (n_sh " +
                                            std::to_string(opt->n_sh) + ")");
    const float* seeds = opt->seeds ? opt->seeds : kDefaultSeeds;
    const int nseeds = opt->seeds ? opt->n_seeds : int(sizeof(kDefaultSeeds) / sizeof(float) / 2);
    if (nseeds < 1 || nseeds > mrcsd::kPeaksMaxSeeds)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_create: 1 to 512 search directions");
    int npeaks = opt->directions ? opt->n_directions : opt->npeaks;
    if (npeaks < 1 || npeaks > mrcsd::kPeaksMaxOut)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_create: 1 to 64 peaks per voxel");
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    auto ctx = std::make_unique<mrpeaks_ctx>();
    MRCSD_HIP(hipGetDevice(&ctx->device));
    ctx->n_sh = opt->n_sh;
    ctx->lmax = lmax;
    ctx->npeaks = npeaks;
    ctx->nseeds = nseeds;
    ctx->threshold = opt->threshold;
    std::vector<float> v(size_t(nseeds) * 3);
    for (int s = 0; s < nseeds; ++s) unit_vector(seeds[2 * s], seeds[2 * s + 1], &v[size_t(s) * 3]);
    ctx->seeds.upload(v.data(), v.size());
    if (opt->directions) {
      std::vector<float> d(size_t(npeaks) * 3);
      for (int i = 0; i < npeaks; ++i) unit_vector(opt->directions[2 * i], opt->directions[2 * i + 1], &d[size_t(i) * 3]);
      ctx->dirs.upload(d.data(), d.size());
      ctx->fixed_dirs = true;
    }
    if (opt->fast) {
      const auto t = fast_table(lmax);
      ctx->fast_table.upload(t.data(), t.size());
    }
    ctx->counter.alloc(1);
    *out = ctx.release();
    return MRCSD_OK;
  });
}

void mrpeaks_destroy(mrpeaks_ctx* ctx) { delete ctx; }

int mrpeaks_npeaks(const mrpeaks_ctx* ctx) { return ctx ? ctx->npeaks : 0; }

int mrpeaks_run(mrpeaks_ctx* ctx, const float* sh, int ld_sh, const uint8_t* mask, const float* peaks_in,
                int ld_peaks_in, size_t nvox, float* out, uint64_t* stats, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!sh || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run: null sh or out");
    if (ld_sh < ctx->n_sh) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run: ld_sh smaller than n_sh");
    if (peaks_in && ctx->fixed_dirs)
      return mrcsd::fail(MRCSD_ERR_ARG, "you can't specify both a peaks file and orientations to be estimated at the same time");
    if (peaks_in && ld_peaks_in < 3 * ctx->npeaks)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run: ld_peaks_in smaller than 3 npeaks");
    MRCSD_HIP(hipSetDevice(ctx->device));
    mrcsd::PeaksLaunch L{};
    L.sh = sh; L.ld_sh = ld_sh; L.n_sh = ctx->n_sh; L.lmax = ctx->lmax; L.nvox = nvox;
    L.seeds = ctx->seeds.get(); L.nseeds = ctx->nseeds; L.npeaks = ctx->npeaks; L.threshold = ctx->threshold;
    L.match = peaks_in ? peaks_in : (ctx->fixed_dirs ? ctx->dirs.get() : nullptr);
    L.fast_table = ctx->fast_table.size() ? ctx->fast_table.get() : nullptr;
    L.ld_match = peaks_in ? ld_peaks_in : 0;
    L.mask = mask; L.out = out;
    L.stats = reinterpret_cast<unsigned long long*>(stats);
    L.counter = ctx->counter.get();
    const size_t items = mrcsd::peaks_work_items(nvox, ctx->nseeds);
    if (ctx->work.size() < items) {
      MRCSD_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));   // a previous launch may use it
      ctx->work.alloc(items);
    }
    L.work = ctx->work.get();
    L.work_items = ctx->work.size();
    MRCSD_HIP(mrcsd::peaks_launch(L, static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

int mrpeaks_run_host(mrpeaks_ctx* ctx, const float* sh, int ld_sh, const uint8_t* mask, const float* peaks_in,
                     int ld_peaks_in, size_t nvox, float* out) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run_host: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!sh || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrpeaks_run_host: null sh or out");
    MRCSD_HIP(hipSetDevice(ctx->device));
    mrcsd::DeviceBuffer<float> d_sh, d_out, d_pk;
    mrcsd::DeviceBuffer<uint8_t> d_mask;
    d_sh.upload(sh, nvox * size_t(ld_sh));
    d_out.alloc(nvox * size_t(3 * ctx->npeaks));
    if (mask) d_mask.upload(mask, nvox);
    if (peaks_in) d_pk.upload(peaks_in, nvox * size_t(ld_peaks_in));
    mrcsd::DeviceBuffer<unsigned long long> d_stats;
    const unsigned long long zero[4] = {0, 0, 0, 0};
    d_stats.upload(zero, 4);
    int rc = mrpeaks_run(ctx, d_sh.get(), ld_sh, mask ? d_mask.get() : nullptr, peaks_in ? d_pk.get() : nullptr,
                         ld_peaks_in, nvox, d_out.get(), reinterpret_cast<uint64_t*>(d_stats.get()), nullptr);
    if (rc) return rc;
    MRCSD_HIP(hipDeviceSynchronize());
    d_out.download(out, nvox * size_t(3 * ctx->npeaks));
    unsigned long long st[4];
    d_stats.download(st, 4);
    if (st[3])
      return mrcsd::fail(MRCSD_ERR_SETUP, "mrpeaks_run_host: " + std::to_string(st[3]) +
                                              " voxel(s) have more than 64 distinct peaks; the reported peaks are "
                                              "chosen among the first 64 (in seed order) only");
    return MRCSD_OK;
  });
}

}  // extern "C"
