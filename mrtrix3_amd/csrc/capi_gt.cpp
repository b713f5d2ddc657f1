// C ABI of global tractography's external energy (include/mrgt.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrgt.h"
#include "../../include/mricls.h"
#include "capi_common.h"
#include "capi_icls_ctx.h"
#include "host/gradient.h"
#include "host/gt.h"
#include "host/icls.h"
#include "kernels/gt.h"

using mrcsd::Matrix;

struct mrgt_ctx {
  int device = 0;
  int nrows = 0, ncols = 0, nf = 0;
  double mu = 0.0;
  Matrix K, Ak;
  mrcsd::DeviceBuffer<double> Kt, Akd;
  mricls_ctx* nnls = nullptr;
  // the fused kernel's tables (nf + 1 <= 4 unknowns): inverses of G's principal submatrices and G
  bool fused = false;
  mrcsd::DeviceBuffer<double> tab;
  // workspace: residuals y [cap][nrows], NNLS solutions [cap][nf + 1], per-workgroup energy sums
  mrcsd::DeviceBuffer<double> y, fk, partial;
  size_t cap = 0;
  ~mrgt_ctx() { mricls_destroy(nnls); }
};

namespace {

// get_DW_scheme, then ExternalEnergyComputer's kernel matrices (externalenergy.cpp:55-101);
// resp_iso [nf][nshell]
void build_kernels(const double* grad, int nvol, const double* resp_wm, int wm_rows, int wm_cols,
                   const double* resp_iso, int nf, int lmax, double weight, Matrix& K, Matrix& Ak) {
  Matrix raw(nvol, 4);
  std::memcpy(raw.data(), grad, sizeof(double) * size_t(nvol) * 4);
  const Matrix g = MR::DWI::get_DW_scheme(raw, nvol);
  Matrix wm(wm_rows, wm_cols);
  std::memcpy(wm.data(), resp_wm, sizeof(double) * size_t(wm_rows) * size_t(wm_cols));
  const size_t nshell = size_t(MR::DWI::Shells(g).count());
  std::vector<std::vector<double>> iso(static_cast<size_t>(nf));
  for (int j = 0; j < nf; ++j) iso[size_t(j)].assign(resp_iso + size_t(j) * nshell, resp_iso + size_t(j + 1) * nshell);
  MR::DWI::Tractography::GT::external_energy_kernels(g, wm, iso, lmax, weight, K, Ak);
}

// The fused kernel's table for n unknowns: for each subset s < 2^n, (G_SS)^-1 embedded in an n x n
// zero matrix, then G = Ak^T Ak.  False when a principal submatrix is not safely invertible (the
// batched ICLS path is taken then).
bool fused_table(const Matrix& Ak, std::vector<double>& tab) {
  const int n = Ak.cols;
  Matrix G(n, n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double a = 0.0;
      for (int r = 0; r < Ak.rows; ++r) a += Ak(r, i) * Ak(r, j);
      G(i, j) = a;
    }
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, G(i, i));
  if (!(dmax > 0.0)) return false;
  tab.assign(size_t((1 << n) * n * n + n * n), 0.0);
  for (int s = 1; s < (1 << n); ++s) {
    std::vector<int> m;
    for (int i = 0; i < n; ++i)
      if ((s >> i) & 1) m.push_back(i);
    const int k = int(m.size());
    // Gauss-Jordan with partial pivoting on [G_SS | I]
    std::vector<double> a(size_t(k * 2 * k), 0.0);
    for (int i = 0; i < k; ++i) {
      for (int j = 0; j < k; ++j) a[size_t(i * 2 * k + j)] = G(m[size_t(i)], m[size_t(j)]);
      a[size_t(i * 2 * k + k + i)] = 1.0;
    }
    for (int c = 0; c < k; ++c) {
      int p = c;
      for (int r = c + 1; r < k; ++r)
        if (std::fabs(a[size_t(r * 2 * k + c)]) > std::fabs(a[size_t(p * 2 * k + c)])) p = r;
      if (!(std::fabs(a[size_t(p * 2 * k + c)]) > 1e-12 * dmax)) return false;
      if (p != c)
        for (int j = 0; j < 2 * k; ++j) std::swap(a[size_t(p * 2 * k + j)], a[size_t(c * 2 * k + j)]);
      const double inv = 1.0 / a[size_t(c * 2 * k + c)];
      for (int j = 0; j < 2 * k; ++j) a[size_t(c * 2 * k + j)] *= inv;
      for (int r = 0; r < k; ++r) {
        if (r == c) continue;
        const double f = a[size_t(r * 2 * k + c)];
        for (int j = 0; j < 2 * k; ++j) a[size_t(r * 2 * k + j)] -= f * a[size_t(c * 2 * k + j)];
      }
    }
    double* Z = tab.data() + size_t(s * n * n);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) Z[m[size_t(i)] * n + m[size_t(j)]] = a[size_t(i * 2 * k + k + j)];
  }
  double* Gt = tab.data() + size_t((1 << n) * n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Gt[i * n + j] = G(i, j);
  return true;
}

}  // namespace

extern "C" {

int mrgt_create(const double* grad, int nvol, const double* resp_wm, int wm_rows, int wm_cols, const double* resp_iso,
                int nf, int lmax, double weight, double ppot, mrgt_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!grad || nvol <= 0 || !resp_wm || wm_rows <= 0 || wm_cols <= 0 || nf < 0 || (nf > 0 && !resp_iso) || !out)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_create: null or empty argument");
    if (lmax < 0 || lmax > 16 || (lmax & 1)) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_create: lmax must be even, 0 to 16");
    if (nf > 63) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_create: at most 63 isotropic tissues");
    *out = nullptr;
    auto ctx = std::make_unique<mrgt_ctx>();
    build_kernels(grad, nvol, resp_wm, wm_rows, wm_cols, resp_iso, nf, lmax, weight, ctx->K, ctx->Ak);
    ctx->nrows = nvol;
    ctx->ncols = ctx->K.cols;
    ctx->nf = nf;
    ctx->mu = ppot * std::sqrt(4.0 * M_PI);   // M_sqrt4PI (externalenergy.cpp:34)
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    MRCSD_HIP(hipGetDevice(&ctx->device));
    // the NNLS problem: Problem(Ak, I) with the defaults (externalenergy.cpp:111)
    Matrix I(nf + 1, nf + 1, 0.0);
    for (int i = 0; i <= nf; ++i) I(i, i) = 1.0;
    MR::Math::ICLS::Problem pr(ctx->Ak, I);
    if ((rc = mrcsd::icls_ctx_from_problem(pr, &ctx->nnls))) return rc;
    const Matrix Kt = ctx->K.transpose();
    ctx->Kt.upload(Kt.data(), Kt.a.size());
    ctx->Akd.upload(ctx->Ak.data(), ctx->Ak.a.size());
    std::vector<double> tab;
    if (mrcsd::gt_fused_fits(nvol, ctx->ncols, nf + 1) && fused_table(ctx->Ak, tab)) {
      ctx->tab.upload(tab.data(), tab.size());
      ctx->fused = true;
    }
    *out = ctx.release();
    return MRCSD_OK;
  });
}

void mrgt_destroy(mrgt_ctx* ctx) { delete ctx; }

int mrgt_kernels(const double* grad, int nvol, const double* resp_wm, int wm_rows, int wm_cols, const double* resp_iso,
                 int nf, int lmax, double weight, double* K, double* Ak) {
  return mrcsd::guarded([&]() -> int {
    if (!grad || nvol <= 0 || !resp_wm || wm_rows <= 0 || wm_cols <= 0 || nf < 0 || (nf > 0 && !resp_iso) || !K || !Ak)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_kernels: null or empty argument");
    if (lmax < 0 || lmax > 16 || (lmax & 1)) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_kernels: lmax must be even, 0 to 16");
    Matrix Km, Akm;
    build_kernels(grad, nvol, resp_wm, wm_rows, wm_cols, resp_iso, nf, lmax, weight, Km, Akm);
    std::memcpy(K, Km.data(), sizeof(double) * Km.a.size());
    std::memcpy(Ak, Akm.data(), sizeof(double) * Akm.a.size());
    return MRCSD_OK;
  });
}

int mrgt_dims(const mrgt_ctx* ctx, int* nrows, int* ncols, int* nf) {
  if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_dims: null context");
  if (nrows) *nrows = ctx->nrows;
  if (ncols) *ncols = ctx->ncols;
  if (nf) *nf = ctx->nf;
  return MRCSD_OK;
}

int mrgt_matrix(const mrgt_ctx* ctx, int which, double* out, size_t capacity) {
  if (!ctx || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_matrix: null argument");
  if (which != 0 && which != 1) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_matrix: unknown matrix id");
  const Matrix& m = which == 0 ? ctx->K : ctx->Ak;
  if (capacity < m.a.size()) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_matrix: buffer too small");
  std::memcpy(out, m.data(), sizeof(double) * m.a.size());
  return MRCSD_OK;
}

int mrgt_energy(mrgt_ctx* ctx, const float* dwi, int ld_dwi, const float* tod, int ld_tod, size_t nvox, float* eext,
                float* fiso, double* total, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx || !total) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy: null context or total");
    if (nvox && (!dwi || !tod || !eext)) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy: null dwi, tod or eext");
    if (ld_dwi < ctx->nrows || ld_tod < ctx->ncols)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy: ld_dwi below the volume count or ld_tod below the TOD size");
    if (nvox > size_t(UINT32_MAX)) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy: more than 2^32 - 1 voxels");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (ctx->fused) {
      const size_t blocks = mrcsd::gt_fused_blocks(nvox);
      if (ctx->partial.size() < blocks + 1) {
        MRCSD_HIP(hipStreamSynchronize(s));   // an earlier call may still use the workspace
        ctx->partial.alloc(blocks + 1);
      }
      MRCSD_HIP(mrcsd::gt_fused(ctx->Kt.get(), ctx->nrows, ctx->ncols, ctx->Akd.get(), ctx->nf + 1, ctx->tab.get(),
                                dwi, ld_dwi, tod, ld_tod, ctx->mu, nvox, eext, fiso, ctx->partial.get(),
                                ctx->partial.size(), total, s));
      return MRCSD_OK;
    }
    if (nvox > ctx->cap) {
      MRCSD_HIP(hipStreamSynchronize(s));   // an earlier call may still use the workspace
      ctx->y.alloc(nvox * size_t(ctx->nrows));
      ctx->fk.alloc(nvox * size_t(ctx->nf + 1));
      ctx->partial.alloc(mrcsd::gt_energy_blocks(nvox) + 1);
      ctx->cap = nvox;
    }
    if (nvox) {
      MRCSD_HIP(mrcsd::gt_residual(ctx->Kt.get(), ctx->nrows, ctx->ncols, dwi, ld_dwi, tod, ld_tod, nvox, ctx->y.get(), s));
      int rc = mrcsd::icls_solve_impl(ctx->nnls, nullptr, ctx->y.get(), ctx->nrows, nullptr, nvox, nullptr,
                                      ctx->fk.get(), nullptr, nullptr, nullptr, nullptr, stream, "mrgt_energy");
      if (rc) return rc;
    }
    if (ctx->partial.size() == 0) ctx->partial.alloc(1);
    MRCSD_HIP(mrcsd::gt_energy(ctx->Akd.get(), ctx->nrows, ctx->nf, ctx->y.get(), ctx->fk.get(), tod, ld_tod, ctx->mu,
                               nvox, eext, fiso, ctx->partial.get(), ctx->partial.size(), total, s));
    return MRCSD_OK;
  });
}

int mrgt_energy_host(mrgt_ctx* ctx, const float* dwi, int ld_dwi, const float* tod, int ld_tod, size_t nvox,
                     float* eext, float* fiso, double* total) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx || !total) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy_host: null context or total");
    if (nvox && (!dwi || !tod || !eext)) return mrcsd::fail(MRCSD_ERR_ARG, "mrgt_energy_host: null dwi, tod or eext");
    MRCSD_HIP(hipSetDevice(ctx->device));
    mrcsd::DeviceBuffer<float> d_dwi, d_tod, d_e, d_f;
    mrcsd::DeviceBuffer<double> d_total;
    d_dwi.upload(dwi, nvox * size_t(ld_dwi));
    d_tod.upload(tod, nvox * size_t(ld_tod));
    d_e.alloc(nvox);
    if (fiso && ctx->nf) d_f.alloc(nvox * size_t(ctx->nf));
    d_total.alloc(1);
    int rc = mrgt_energy(ctx, d_dwi.get(), ld_dwi, d_tod.get(), ld_tod, nvox, d_e.get(),
                         (fiso && ctx->nf) ? d_f.get() : nullptr, d_total.get(), nullptr);
    if (rc) return rc;
    MRCSD_HIP(hipDeviceSynchronize());
    d_e.download(eext, nvox);
    if (fiso && ctx->nf) d_f.download(fiso, nvox * size_t(ctx->nf));
    d_total.download(total, 1);
    return MRCSD_OK;
  });
}

}  // extern "C"
