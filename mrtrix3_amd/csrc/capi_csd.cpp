// C ABI of the CSD path (include/mrcsd.h): host-side CSD::Shared set-up and the GPU solver
// context.  No C++ exception crosses the boundary; errors become a status code plus a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mrcsd.h"
#include "capi_common.h"
#include "capi_csd_ctx.h"
#include "host/csd.h"
#include "host/gradient.h"
#include "kernels/csd_solve.h"

using MR::DWI::SDeconv::CSD;
using MR::DWI::SDeconv::CSDOptions;
using mrcsd::Matrix;

struct mrcsd_shared {
  std::unique_ptr<CSD::Shared> shared;
};

namespace {

// Cholesky of the symmetric n x n matrix A; true when every pivot exceeds `rel` times the largest
// diagonal entry (positive definite with margin above rounding).
bool positive_definite(const double* A, int n, double rel) {
  std::vector<double> L(size_t(n) * n, 0.0);
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, A[size_t(i) * n + i]);
  if (!(dmax > 0.0)) return false;
  for (int j = 0; j < n; ++j) {
    double d = A[size_t(j) * n + j];
    for (int k = 0; k < j; ++k) d -= L[size_t(j) * n + k] * L[size_t(j) * n + k];
    if (!(d > rel * dmax)) return false;
    const double ljj = std::sqrt(d);
    L[size_t(j) * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = A[size_t(i) * n + j];
      for (int k = 0; k < j; ++k) s -= L[size_t(i) * n + k] * L[size_t(j) * n + k];
      L[size_t(i) * n + j] = s / ljj;
    }
  }
  return true;
}

}  // namespace

extern "C" {

void mrcsd_options_default(mrcsd_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->neg_lambda = MR::DWI::SDeconv::DEFAULT_CSD_NEG_LAMBDA;
  opt->norm_lambda = MR::DWI::SDeconv::DEFAULT_CSD_NORM_LAMBDA;
  opt->threshold = MR::DWI::SDeconv::DEFAULT_CSD_THRESHOLD;
  opt->niter = int(MR::DWI::SDeconv::DEFAULT_CSD_NITER);
}

int mrcsd_shared_create(const double* grad, int nvol, const double* response, int n_resp,
                        const mrcsd_options* opt, mrcsd_shared** out) {
  return mrcsd::guarded([&]() -> int {
    if (!grad || nvol <= 0 || !response || n_resp <= 0 || !out)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_shared_create: null or empty argument");
    mrcsd_options o;
    if (opt) o = *opt; else mrcsd_options_default(&o);
    Matrix raw(nvol, 4);
    std::memcpy(raw.data(), grad, sizeof(double) * size_t(nvol) * 4);
    const auto scaling = o.bvalue_scaling == 1 ? MR::DWI::BValueScaling::UserOn
                         : o.bvalue_scaling == 2 ? MR::DWI::BValueScaling::UserOff
                                                 : MR::DWI::BValueScaling::Auto;
    Matrix g = MR::DWI::get_DW_scheme(raw, nvol, scaling);
    std::vector<double> requested;
    if (o.shells && o.n_shells > 0) requested.assign(o.shells, o.shells + o.n_shells);
    auto s = std::make_unique<CSD::Shared>(g, requested);
    CSDOptions co;
    if (o.lmax) co.lmax.push_back(uint32_t(o.lmax));
    if (o.filter && o.n_filter > 0) co.filter.assign(o.filter, o.filter + o.n_filter);
    if (o.directions && o.n_directions > 0) {
      co.directions = Matrix(o.n_directions, 2);
      std::memcpy(co.directions.data(), o.directions, sizeof(double) * size_t(o.n_directions) * 2);
    }
    co.has_neg_lambda = true; co.neg_lambda = o.neg_lambda;
    co.has_norm_lambda = true; co.norm_lambda = o.norm_lambda;
    co.has_threshold = true; co.threshold = o.threshold;
    co.has_niter = true; co.niter = o.niter;
    s->parse_cmdline_options(co);
    s->set_response(std::vector<double>(response, response + n_resp));
    s->init();
    auto* h = new mrcsd_shared;
    h->shared = std::move(s);
    *out = h;
    return MRCSD_OK;
  });
}

void mrcsd_shared_destroy(mrcsd_shared* s) { delete s; }

int mrcsd_shared_info_get(const mrcsd_shared* s, mrcsd_shared_info* info) {
  if (!s || !info) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_shared_info_get: null argument");
  const CSD::Shared& S = *s->shared;
  info->n_dw = int(S.dwis.size());
  info->n_sh = int(S.nSH());
  info->n_init = S.rconv.rows;
  info->n_hr = S.HR_trans.rows;
  info->lmax = int(S.lmax);
  info->lmax_response = int(S.lmax_response);
  info->lmax_data = int(S.lmax_data);
  info->niter = int(S.niter);
  info->threshold = S.threshold;
  info->norm_lambda = S.norm_lambda;
  return MRCSD_OK;
}

int mrcsd_shared_matrix(const mrcsd_shared* s, int which, double* out, size_t capacity) {
  if (!s || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_shared_matrix: null argument");
  const CSD::Shared& S = *s->shared;
  std::vector<double> v;
  switch (which) {
    case 0: v = S.rconv.a; break;
    case 1: v = S.M.a; break;
    case 2: v = S.Mt_M.a; break;
    case 3: v = S.HR_trans.a; break;
    case 4: v.assign(S.dwis.begin(), S.dwis.end()); break;
    case 5: v = S.grad.a; break;
    default: return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_shared_matrix: unknown matrix id");
  }
  if (capacity < v.size()) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_shared_matrix: buffer too small");
  std::memcpy(out, v.data(), sizeof(double) * v.size());
  return MRCSD_OK;
}

int mrcsd_create(const mrcsd_params* p, mrcsd_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!p || !out || !p->rconv || !p->M || !p->MtM || !p->HR_trans || !p->dwis)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: null argument");
    if (!mrcsd::csd_supported_nsh(p->n_sh))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: n_sh = " + std::to_string(p->n_sh) +
                                            " not supported (lmax 2 to 12)");
    if (p->n_dw <= 0 || p->n_dw > 256) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: n_dw must be in [1, 256]");
    if (p->n_hr <= 0 || p->n_hr > 512) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: n_hr must be in [1, 512]");
    if (p->n_init <= 0 || p->n_init > p->n_sh) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: bad n_init");
    if (p->niter < 0 || p->niter > 65534) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: bad niter");
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    auto ctx = std::make_unique<mrcsd_ctx>();
    MRCSD_HIP(hipGetDevice(&ctx->device));
    const int n = p->n_sh, R = p->n_hr, N = p->n_dw, ni = p->n_init;
    const int NT = mrcsd::csd_threads_for(n);
    const int Rpad = (R + NT - 1) / NT * NT;
    std::vector<double> HRt(size_t(n) * Rpad, 0.0), rconvT(size_t(N) * ni);
    for (int r = 0; r < R; ++r)
      for (int j = 0; j < n; ++j) HRt[(size_t(r / NT) * n + j) * NT + r % NT] = p->HR_trans[size_t(r) * n + j];
    for (int i = 0; i < ni; ++i)
      for (int c = 0; c < N; ++c) rconvT[size_t(c) * ni + i] = p->rconv[size_t(i) * N + c];
    ctx->MtM.upload(p->MtM, size_t(n) * n);
    ctx->HR.upload(p->HR_trans, size_t(R) * n);
    ctx->HRt.upload(HRt.data(), HRt.size());
    ctx->M.upload(p->M, size_t(N) * n);
    ctx->rconvT.upload(rconvT.data(), rconvT.size());
    for (int i = 0; i < N; ++i) {
      if (p->dwis[i] < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create: negative volume index");
      ctx->max_vol = std::max(ctx->max_vol, p->dwis[i]);
    }
    ctx->dwis.upload(p->dwis, size_t(N));
    ctx->counter.alloc(1);
    mrcsd::CsdDeviceProblem& P = ctx->P;
    P.MtM = ctx->MtM.get(); P.HR = ctx->HR.get(); P.HRt = ctx->HRt.get();
    P.M = ctx->M.get(); P.rconvT = ctx->rconvT.get(); P.dwis = ctx->dwis.get();
    P.thr = p->threshold;
    P.n_dw = N; P.n_sh = n; P.n_init = ni; P.n_hr = R; P.n_hr_pad = Rpad; P.niter = p->niter;
    ctx->warm_certifiable = p->threshold == 0.0 && positive_definite(p->MtM, n, 1e-12);
    *out = ctx.release();
    return MRCSD_OK;
  });
}

int mrcsd_create_from_shared(const mrcsd_shared* s, mrcsd_ctx** out) {
  if (!s) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_create_from_shared: null shared");
  const CSD::Shared& S = *s->shared;
  std::vector<int> dwis(S.dwis.begin(), S.dwis.end());
  mrcsd_params p;
  p.n_dw = int(S.dwis.size());
  p.n_sh = int(S.nSH());
  p.n_init = S.rconv.rows;
  p.n_hr = S.HR_trans.rows;
  p.rconv = S.rconv.data();
  p.M = S.M.data();
  p.MtM = S.Mt_M.data();
  p.HR_trans = S.HR_trans.data();
  p.dwis = dwis.data();
  p.threshold = S.threshold;
  p.niter = int(S.niter);
  return mrcsd_create(&p, out);
}

void mrcsd_destroy(mrcsd_ctx* ctx) { delete ctx; }

int mrcsd_warm_supported(const mrcsd_ctx* ctx) { return ctx && ctx->warm_certifiable ? 1 : 0; }

int mrcsd_run(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
              float* fod, const mrcsd_outputs* out, const mrcsd_warm* warm, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!dw || !fod) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run: null dw or fod");
    if (ld_dw <= 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run: ld_dw must be positive");
    const int maxvol = ctx->max_vol;
    if (maxvol >= ld_dw)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run: ld_dw = " + std::to_string(ld_dw) +
                                            " does not cover volume index " + std::to_string(maxvol));
    MRCSD_HIP(hipSetDevice(ctx->device));
    mrcsd::CsdLaunch L{};
    L.dw = dw; L.ld_dw = ld_dw; L.valid = valid; L.nvox = nvox; L.fod = fod;
    if (out) {
      L.iters = out->iters;
      L.nneg = out->nneg;
      L.negmask = reinterpret_cast<unsigned long long*>(out->negmask);
      L.stats = reinterpret_cast<unsigned long long*>(out->stats);
    }
    L.warm_F0 = warm ? warm->F0 : nullptr;
    L.warm_neg = warm ? reinterpret_cast<const unsigned long long*>(warm->neg) : nullptr;
    L.warm_budget = warm ? warm->budget : 0;
    if (warm && (warm->F0 || warm->neg) && (warm->budget < 1 || warm->budget > 65534))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run: warm-start budget must be in [1, 65534]");
    if (warm && (warm->F0 || warm->neg) && !ctx->warm_certifiable)
      return mrcsd::fail(MRCSD_ERR_ARG,
                         "mrcsd_run: a warm start needs threshold 0 and a positive definite normal matrix "
                         "(norm_lambda > 0, or M of full column rank); solve without a proposal");
    L.counter = ctx->counter.get();
    L.max_groups = ctx->groups[(L.warm_F0 || L.warm_neg) ? 1 : 0];
    MRCSD_HIP(mrcsd::csd_launch(ctx->P, L, static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

int mrcsd_run_host(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
                   float* fod, uint16_t* iters, uint16_t* nneg) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run_host: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!dw || !fod) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_run_host: null dw or fod");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const size_t n = size_t(ctx->P.n_sh);
    mrcsd::DeviceBuffer<float> d_dw, d_fod;
    mrcsd::DeviceBuffer<uint8_t> d_valid;
    mrcsd::DeviceBuffer<uint16_t> d_it, d_nn;
    d_dw.upload(dw, nvox * size_t(ld_dw));
    d_fod.alloc(nvox * n);
    if (valid) d_valid.upload(valid, nvox);
    if (iters) d_it.alloc(nvox);
    if (nneg) d_nn.alloc(nvox);
    mrcsd_outputs o{};
    o.iters = iters ? d_it.get() : nullptr;
    o.nneg = nneg ? d_nn.get() : nullptr;
    int rc = mrcsd_run(ctx, d_dw.get(), ld_dw, valid ? d_valid.get() : nullptr, nvox, d_fod.get(), &o,
                       nullptr, nullptr);
    if (rc) return rc;
    MRCSD_HIP(hipDeviceSynchronize());
    d_fod.download(fod, nvox * n);
    if (iters) d_it.download(iters, nvox);
    if (nneg) d_nn.download(nneg, nvox);
    return MRCSD_OK;
  });
}

}  // extern "C"
