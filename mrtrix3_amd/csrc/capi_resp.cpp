// C ABI of amp2response (include/mrresp.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/mricls.h"
#include "../../include/mrresp.h"
#include "capi_common.h"
#include "host/dense.h"
#include "host/sh.h"
#include "kernels/resp.h"

struct mrresp_ctx {
  int device = 0;
  int lmax = 0, nvol = 0, nblocks = 0;
  mrcsd::DeviceBuffer<double> dirs, inv_norm, partial, out;
  mrcsd::DeviceBuffer<int> vols;
  int max_vol = 0;   // largest volume index (checked against each call's row length)
};

namespace {

constexpr int kRespBlocks = 1024;   // fixed, so sums are formed in the same order on any device

// M rf = b by Cholesky (amp2response.cpp:339: M.llt().solve(b), M's lower triangle)
bool llt_solve(int n, const double* M, const double* b, double* x) {
  mrcsd::Matrix L(n, n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) L(r, c) = M[r * n + c];
  if (!mrcsd::cholesky_lower(L)) return false;
  std::vector<double> y(size_t(n), 0.0);
  for (int r = 0; r < n; ++r) {
    double s = b[r];
    for (int k = 0; k < r; ++k) s -= L(r, k) * y[size_t(k)];
    y[size_t(r)] = s / L(r, r);
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = y[size_t(r)];
    for (int k = r + 1; k < n; ++k) s -= L(k, r) * x[k];
    x[r] = s / L(r, r);
  }
  return true;
}

}  // namespace

extern "C" {

int mrresp_create(const double* dirs, const int* volumes, int nvol, int lmax, mrresp_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!dirs || !volumes || !out || nvol < 1) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_create: null or empty argument");
    if (lmax < 0 || lmax > mrcsd::kRespMaxLmax || lmax % 2)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_create: lmax must be even and in [0, 12]");
    for (int i = 0; i < nvol; ++i) {
      if (volumes[i] < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_create: negative volume index");
      if (lmax && (!std::isfinite(dirs[3 * i]) || !std::isfinite(dirs[3 * i + 1]) || !std::isfinite(dirs[3 * i + 2])))
        return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_create: non-finite direction (at lmax 0 the directions do not enter)");
    }
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    auto ctx = std::make_unique<mrresp_ctx>();
    MRCSD_HIP(hipGetDevice(&ctx->device));
    ctx->lmax = lmax;
    ctx->nvol = nvol;
    ctx->nblocks = kRespBlocks;
    ctx->dirs.upload(dirs, size_t(nvol) * 3);
    std::vector<double> inv(static_cast<size_t>(nvol));
    for (int i = 0; i < nvol; ++i) {
      const double* g = dirs + 3 * i;
      inv[size_t(i)] = 1.0 / std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    }
    ctx->inv_norm.upload(inv.data(), inv.size());
    ctx->vols.upload(volumes, size_t(nvol));
    for (int i = 0; i < nvol; ++i) ctx->max_vol = std::max(ctx->max_vol, volumes[i]);
    ctx->partial.alloc(size_t(kRespBlocks) * size_t(mrcsd::resp_values(lmax)));
    ctx->out.alloc(size_t(mrcsd::resp_values(lmax)));
    *out = ctx.release();
    return MRCSD_OK;
  });
}

void mrresp_destroy(mrresp_ctx* ctx) { delete ctx; }

int mrresp_ncoef(const mrresp_ctx* ctx) { return ctx ? mrcsd::resp_nz(ctx->lmax) : 0; }

int mrresp_accumulate(mrresp_ctx* ctx, const float* amp, int ld_amp, const float* fibre, int ld_fibre,
                      const uint8_t* mask, size_t nvox, double* M, double* b, uint64_t* count, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx || !M || !b) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_accumulate: null argument");
    if (nvox && (!amp || !fibre)) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_accumulate: null input");
    if (ld_fibre < 3) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_accumulate: the fibre rows need 3 values");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const int nz = mrcsd::resp_nz(ctx->lmax), nv = mrcsd::resp_values(ctx->lmax);
    if (ctx->max_vol >= ld_amp) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_accumulate: a volume index is past the row length");
    std::vector<double> host(size_t(nv), 0.0);
    if (nvox) {
      mrcsd::RespLaunch L{};
      L.amp = amp;
      L.ld_amp = ld_amp;
      L.fdir = fibre;
      L.ld_fdir = ld_fibre;
      L.mask = mask;
      L.nvox = nvox;
      L.dirs = ctx->dirs.get();
      L.inv_norm = ctx->inv_norm.get();
      L.vols = ctx->vols.get();
      L.nvol = ctx->nvol;
      L.lmax = ctx->lmax;
      L.partial = ctx->partial.get();
      L.out = ctx->out.get();
      L.nblocks = ctx->nblocks;
      hipStream_t s = static_cast<hipStream_t>(stream);
      MRCSD_HIP(mrcsd::resp_accumulate_launch(L, s));
      MRCSD_HIP(hipMemcpyAsync(host.data(), ctx->out.get(), host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      MRCSD_HIP(hipStreamSynchronize(s));
    }
    int q = 0;
    for (int r = 0; r < nz; ++r)
      for (int c = 0; c <= r; ++c, ++q) M[r * nz + c] = M[c * nz + r] = host[size_t(q)];
    for (int r = 0; r < nz; ++r) b[r] = host[size_t(q + r)];
    if (count) *count = (uint64_t)host[size_t(q + nz)];
    return MRCSD_OK;
  });
}

int mrresp_solve(int lmax, const double* M, const double* b, int constrained, double* rf, int* iterations) {
  return mrcsd::guarded([&]() -> int {
    if (!M || !b || !rf) return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_solve: null argument");
    if (lmax < 0 || lmax > mrcsd::kRespMaxLmax || lmax % 2)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrresp_solve: lmax must be even and in [0, 12]");
    const int n = mrcsd::resp_nz(lmax);
    if (iterations) *iterations = 0;
    if (!lmax || !constrained) {
      if (!llt_solve(n, M, b, rf)) return mrcsd::fail(MRCSD_ERR_SETUP, "mrresp_solve: M is not positive definite");
      return MRCSD_OK;
    }
    // constraints: amplitudes and elevation derivatives at 0, 1, ..., 90 degrees
    // (amp2response.cpp:343-354, ZSH.h:66-99)
    constexpr int kAngles = 91;
    std::vector<double> A(size_t(2 * kAngles) * size_t(n), 0.0);
    std::vector<double> P(size_t(lmax) + 2);
    for (int i = 0; i < kAngles; ++i) {
      const double el = double(i) * M_PI / 180.0;
      const double x = std::cos(el);
      mrcsd::sh::legendre_column(P.data(), lmax, 0, x);
      for (int l = 0; l <= lmax; l += 2) A[size_t(i) * n + size_t(l / 2)] = P[size_t(l)];
      mrcsd::sh::legendre_column(P.data(), lmax, 1, x);
      for (int l = 2; l <= lmax; l += 2)
        A[size_t(kAngles + i) * n + size_t(l / 2)] = P[size_t(l)] * std::sqrt(double(l * (l + 1)));
    }
    mricls_ctx* icls = nullptr;
    int rc = mricls_create(M, n, n, A.data(), 2 * kAngles, nullptr, 0, 1e-10, 1e-10, 0, 0.0, 1, &icls);
    if (rc) return rc;
    // b and the response in double, as Problem<default_type> takes and returns them (:357-361)
    uint16_t it = 0, na = 0;
    rc = mricls_solve_host_f64(icls, b, n, nullptr, 1, rf, &it, &na);
    mricls_destroy(icls);
    if (rc) return rc;
    if (iterations) *iterations = int(it);
    return MRCSD_OK;
  });
}

}  // extern "C"
