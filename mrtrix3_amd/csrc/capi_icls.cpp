// C ABI of the ICLS / MSMT-CSD path (include/mricls.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mricls.h"
#include "capi_common.h"
#include "capi_icls_ctx.h"
#include "host/gradient.h"
#include "host/icls.h"
#include "host/msmt.h"
#include "kernels/icls_solve.h"

using MR::Math::ICLS::Problem;
using mrcsd::Matrix;

struct mrmsmt_shared {
  std::unique_ptr<MR::DWI::SDeconv::MSMT_CSD::Shared> shared;
};

namespace mrcsd {

int icls_ctx_from_problem(const Problem& pr, mricls_ctx** out) {
  const int n = int(pr.num_parameters()), nm = pr.b2d.rows, ncon = pr.B.rows;
  const int num_eq = int(pr.num_eq);
  if (n < 1 || n > mrcsd::kIclsMaxParams)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls: number of parameters must be in [1, 128]");
  if (nm < 1)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls: no measurements");
  if (ncon < 1 || ncon > mrcsd::kIclsMaxCon)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls: number of constraints must be in [1, 1024]");
  if (num_eq > mrcsd::kIclsMaxActive)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls: at most 128 equality constraints");
  if (pr.max_niter > 65533) return mrcsd::fail(MRCSD_ERR_ARG, "mricls: max_iter too large");
  int rc = mrcsd::require_gpu();
  if (rc) return rc;
  const int ldg = 64 * mrcsd::icls_constraint_words(ncon);
  std::vector<double> Bt(size_t(n) * ldg, 0.0), t(size_t(ldg), 0.0), G(size_t(ncon) * ldg, 0.0);
  for (int r = 0; r < ncon; ++r)
    for (int j = 0; j < n; ++j) Bt[size_t(j) * ldg + r] = pr.B(r, j);
  for (int r = 0; r < ncon && !pr.t.empty(); ++r) t[size_t(r)] = pr.t[size_t(r)];
  for (int a = 0; a < ncon; ++a)
    for (int b = 0; b <= a; ++b) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += pr.B(a, j) * pr.B(b, j);
      G[size_t(a) * ldg + b] = s;
      G[size_t(b) * ldg + a] = s;
    }
  std::vector<double> Lf(size_t(n) * n, 0.0);   // the lower factor, upper triangle zero
  for (int r = 0; r < n; ++r)
    for (int c = 0; c <= r; ++c) Lf[size_t(r) * n + c] = pr.chol_HtH(r, c);

  auto ctx = std::make_unique<mricls_ctx>();
  MRCSD_HIP(hipGetDevice(&ctx->device));
  ctx->b2d.upload(pr.b2d.data(), size_t(nm) * n);
  ctx->Bt.upload(Bt.data(), Bt.size());
  ctx->t.upload(t.data(), t.size());
  ctx->G.upload(G.data(), G.size());
  ctx->L.upload(Lf.data(), Lf.size());
  ctx->standard_form = pr.standard_form;
  if (!pr.standard_form) {
    std::vector<double> Ht(size_t(n) * nm);
    for (int r = 0; r < nm; ++r)
      for (int j = 0; j < n; ++j) Ht[size_t(j) * nm + r] = pr.H(r, j);
    ctx->Ht.upload(Ht.data(), Ht.size());
  }
  ctx->counters.alloc(2 * mrcsd::kIclsMaxPasses);
  ctx->capacity = mrcsd::icls_default_capacity(n, num_eq);
  mrcsd::IclsDeviceProblem& P = ctx->P;
  P.b2d = ctx->b2d.get(); P.Bt = ctx->Bt.get(); P.t = ctx->t.get(); P.G = ctx->G.get();
  P.L = ctx->L.get();
  P.lambda_min = pr.lambda_min_norm;
  P.tol = pr.tol;
  P.nmeas = nm; P.n = n; P.ncon = ncon; P.ncon_pad = ldg; P.num_eq = num_eq;
  P.max_niter = int(pr.max_niter);
  *out = ctx.release();
  return MRCSD_OK;
}

int icls_solve_impl(mricls_ctx* ctx, const float* b, const double* b64, int ld_b, const uint8_t* valid,
                    size_t nvox, float* x, double* x64, uint16_t* iters, uint16_t* nactive, uint64_t* stats,
                    const IclsScratch* scratch, void* stream, const char* who) {
  const std::string fn(who);
  if (!ctx) return fail(MRCSD_ERR_ARG, fn + ": null context");
  if (nvox == 0) return MRCSD_OK;
  if ((!b && !b64) || (!x && !x64)) return fail(MRCSD_ERR_ARG, fn + ": null b or x");
  if (ld_b < ctx->P.nmeas) return fail(MRCSD_ERR_ARG, fn + ": ld_b smaller than nmeas");
  if (nvox > size_t(UINT32_MAX)) return fail(MRCSD_ERR_ARG, fn + ": more than 2^32 - 1 voxels");
  MRCSD_HIP(hipSetDevice(ctx->device));
  IclsScratch S{};
  if (scratch) {
    S = *scratch;
  } else {
    if (ctx->capacity < kIclsMaxActive && ctx->queue_cap < nvox) {
      // queues of voxels passed on to larger capacities: grown (synchronously) to this call's size
      MRCSD_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
      ctx->queue.alloc(2 * nvox);
      ctx->queue_cap = nvox;
    }
    S.counters = ctx->counters.get();
    S.queue[0] = ctx->queue.get();
    S.queue[1] = ctx->queue.get() ? ctx->queue.get() + ctx->queue_cap : nullptr;
    S.capacity = ctx->queue_cap;
  }
  IclsLaunch L{};
  L.b = b; L.b64 = b64; L.ld_b = ld_b; L.valid = valid; L.nvox = nvox; L.x = x; L.x64 = x64; L.iters = iters;
  L.nactive = nactive; L.stats = reinterpret_cast<unsigned long long*>(stats);
  MRCSD_HIP(icls_launch(ctx->P, L, S, ctx->capacity, static_cast<hipStream_t>(stream)));
  return MRCSD_OK;
}

}  // namespace mrcsd

namespace {

int make_ctx(const Problem& pr, mricls_ctx** out) { return mrcsd::icls_ctx_from_problem(pr, out); }

// mricls_solve_host and its double-precision twin: copies in, solves, copies out.
template <class T>
int solve_host(mricls_ctx* ctx, const T* b, int ld_b, const uint8_t* valid, size_t nvox, T* x, uint16_t* iters,
               uint16_t* nactive, const char* who) {
  const std::string fn(who);
  if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, fn + ": null context");
  if (nvox == 0) return MRCSD_OK;
  if (!b || !x) return mrcsd::fail(MRCSD_ERR_ARG, fn + ": null b or x");
  MRCSD_HIP(hipSetDevice(ctx->device));
  const size_t n = size_t(ctx->P.n);
  mrcsd::DeviceBuffer<T> d_b, d_x;
  mrcsd::DeviceBuffer<uint8_t> d_v;
  mrcsd::DeviceBuffer<uint16_t> d_it, d_na;
  d_b.upload(b, nvox * size_t(ld_b));
  d_x.alloc(nvox * n);
  if (valid) d_v.upload(valid, nvox);
  d_it.alloc(nvox);
  if (nactive) d_na.alloc(nvox);
  constexpr bool f64 = std::is_same<T, double>::value;
  int rc = mrcsd::icls_solve_impl(ctx, f64 ? nullptr : reinterpret_cast<const float*>(d_b.get()),
                                  f64 ? reinterpret_cast<const double*>(d_b.get()) : nullptr, ld_b,
                                  valid ? d_v.get() : nullptr, nvox,
                                  f64 ? nullptr : reinterpret_cast<float*>(d_x.get()),
                                  f64 ? reinterpret_cast<double*>(d_x.get()) : nullptr, d_it.get(),
                                  nactive ? d_na.get() : nullptr, nullptr, nullptr, nullptr, who);
  if (rc) return rc;
  MRCSD_HIP(hipDeviceSynchronize());
  std::vector<uint16_t> it(nvox);
  d_it.download(it.data(), nvox);
  d_x.download(x, nvox * n);
  if (iters) std::memcpy(iters, it.data(), nvox * sizeof(uint16_t));
  if (nactive) d_na.download(nactive, nvox);
  const size_t over = size_t(std::count(it.begin(), it.end(), uint16_t(0xFFFE)));
  if (over)
    return mrcsd::fail(MRCSD_ERR_SETUP, fn + ": " + std::to_string(over) +
                                            " voxel(s) need more than 128 active constraints (iters 0xFFFE)");
  return MRCSD_OK;
}

}  // namespace

extern "C" {

int mricls_create(const double* H, int nmeas, int n, const double* A, int ncon, const double* t,
                  int num_eq, double sol_reg, double con_reg, int max_iter, double tol,
                  int standard_form, mricls_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!H || !A || !out || nmeas <= 0 || n <= 0 || ncon <= 0)
      return mrcsd::fail(MRCSD_ERR_ARG, "mricls_create: null or empty argument");
    if (num_eq < 0 || num_eq > ncon) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_create: bad num_eq");
    if (max_iter < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_create: negative max_iter");
    if (standard_form && nmeas != n)
      return mrcsd::fail(MRCSD_ERR_ARG, "mricls_create: standard form needs a square [n][n] matrix");
    Matrix Hm(nmeas, n), Am(ncon, n);
    std::memcpy(Hm.data(), H, sizeof(double) * size_t(nmeas) * n);
    std::memcpy(Am.data(), A, sizeof(double) * size_t(ncon) * n);
    std::vector<double> tv;
    if (t) tv.assign(t, t + ncon);
    Problem pr(Hm, Am, tv, size_t(num_eq), sol_reg, con_reg, size_t(max_iter), tol, standard_form != 0);
    return make_ctx(pr, out);
  });
}

void mricls_destroy(mricls_ctx* ctx) { delete ctx; }

int mricls_dims(const mricls_ctx* ctx, int* nmeas, int* n, int* ncon, int* max_niter) {
  if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_dims: null context");
  if (nmeas) *nmeas = ctx->P.nmeas;
  if (n) *n = ctx->P.n;
  if (ncon) *ncon = ctx->P.ncon;
  if (max_niter) *max_niter = ctx->P.max_niter;
  return MRCSD_OK;
}

int mricls_set_capacity(mricls_ctx* ctx, int slots) {
  if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_set_capacity: null context");
  if (slots == 0) {
    ctx->capacity = mrcsd::icls_default_capacity(ctx->P.n, ctx->P.num_eq);
    return MRCSD_OK;
  }
  if (!mrcsd::icls_valid_capacity(slots))
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls_set_capacity: slots must be 0, 16, 32, 48, 64, 96 or 128");
  if (slots < ctx->P.num_eq)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls_set_capacity: fewer slots than equality constraints");
  ctx->capacity = slots;
  return MRCSD_OK;
}

int mricls_get_capacity(const mricls_ctx* ctx) { return ctx ? ctx->capacity : 0; }

int mricls_solve(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                 float* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream) {
  return mrcsd::guarded([&]() -> int {
    return mrcsd::icls_solve_impl(ctx, b, nullptr, ld_b, valid, nvox, x, nullptr, iters, nactive, stats, nullptr,
                                  stream, "mricls_solve");
  });
}

int mricls_solve_f64(mricls_ctx* ctx, const double* b, int ld_b, const uint8_t* valid, size_t nvox,
                     double* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream) {
  return mrcsd::guarded([&]() -> int {
    return mrcsd::icls_solve_impl(ctx, nullptr, b, ld_b, valid, nvox, nullptr, x, iters, nactive, stats, nullptr,
                                  stream, "mricls_solve_f64");
  });
}

int mricls_predict(mricls_ctx* ctx, const float* x, int ld_x, size_t nvox, float* out, int ld_out, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_predict: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!x || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_predict: null x or out");
    if (ctx->standard_form)
      return mrcsd::fail(MRCSD_ERR_ARG, "mricls_predict: the problem was given in standard form (no problem matrix)");
    if (ld_x < ctx->P.n || ld_out < ctx->P.nmeas)
      return mrcsd::fail(MRCSD_ERR_ARG, "mricls_predict: ld_x below n or ld_out below nmeas");
    MRCSD_HIP(hipSetDevice(ctx->device));
    MRCSD_HIP(mrcsd::icls_predict(ctx->Ht.get(), ctx->P.n, ctx->P.nmeas, x, ld_x, nvox, out, ld_out,
                                  static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

int mricls_solve_host(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                      float* x, uint16_t* iters, uint16_t* nactive) {
  return mrcsd::guarded([&]() -> int {
    return solve_host<float>(ctx, b, ld_b, valid, nvox, x, iters, nactive, "mricls_solve_host");
  });
}

int mricls_solve_host_f64(mricls_ctx* ctx, const double* b, int ld_b, const uint8_t* valid, size_t nvox,
                          double* x, uint16_t* iters, uint16_t* nactive) {
  return mrcsd::guarded([&]() -> int {
    return solve_host<double>(ctx, b, ld_b, valid, nvox, x, iters, nactive, "mricls_solve_host_f64");
  });
}

int mrmsmt_shared_create(const double* grad, int nvol, const double* responses,
                         const int* resp_rows, const int* resp_cols, int ntissue, const int* lmax,
                         int n_lmax, const double* directions, int n_directions,
                         double norm_lambda, double neg_lambda, const double* shells, int n_shells,
                         mrmsmt_shared** out) {
  return mrcsd::guarded([&]() -> int {
    if (!grad || nvol <= 0 || !responses || !resp_rows || !resp_cols || ntissue <= 0 || !out)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_create: null or empty argument");
    Matrix raw(nvol, 4);
    std::memcpy(raw.data(), grad, sizeof(double) * size_t(nvol) * 4);
    Matrix g = MR::DWI::get_DW_scheme(raw, nvol);
    std::vector<double> requested;
    if (shells && n_shells > 0) requested.assign(shells, shells + n_shells);
    auto s = std::make_unique<MR::DWI::SDeconv::MSMT_CSD::Shared>(g, requested);
    if (lmax && n_lmax > 0) {
      for (int i = 0; i < n_lmax; ++i)
        if (lmax[i] < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_create: negative lmax");
      s->lmax.assign(lmax, lmax + n_lmax);
    }
    if (directions && n_directions > 0) {
      s->HR_dirs = Matrix(n_directions, 2);
      std::memcpy(s->HR_dirs.data(), directions, sizeof(double) * size_t(n_directions) * 2);
    }
    if (norm_lambda >= 0.0) s->solution_min_norm_regularisation = norm_lambda;
    if (neg_lambda >= 0.0) s->constraint_min_norm_regularisation = neg_lambda;
    std::vector<Matrix> resp;
    size_t off = 0;
    for (int t = 0; t < ntissue; ++t) {
      Matrix r(resp_rows[t], resp_cols[t]);
      std::memcpy(r.data(), responses + off, sizeof(double) * size_t(resp_rows[t]) * resp_cols[t]);
      off += size_t(resp_rows[t]) * resp_cols[t];
      resp.push_back(r);
    }
    s->set_responses(resp);
    s->init();
    auto* h = new mrmsmt_shared;
    h->shared = std::move(s);
    *out = h;
    return MRCSD_OK;
  });
}

void mrmsmt_shared_destroy(mrmsmt_shared* s) { delete s; }

int mrmsmt_shared_info(const mrmsmt_shared* s, int* nshell, int* nparams, int* ncon, int* lmax_out) {
  if (!s) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_info: null argument");
  const auto& S = *s->shared;
  if (nshell) *nshell = int(S.num_shells());
  if (nparams) *nparams = S.C.cols;
  if (ncon) *ncon = S.A.rows;
  if (lmax_out)
    for (size_t t = 0; t < S.lmax.size(); ++t) lmax_out[t] = int(S.lmax[t]);
  return MRCSD_OK;
}

int mrmsmt_shared_matrix(const mrmsmt_shared* s, int which, double* out, size_t capacity) {
  if (!s || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_matrix: null argument");
  const Matrix& m = which == 0 ? s->shared->C : s->shared->A;
  if (which != 0 && which != 1) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_matrix: unknown id");
  if (capacity < m.a.size()) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_shared_matrix: buffer too small");
  std::memcpy(out, m.data(), sizeof(double) * m.a.size());
  return MRCSD_OK;
}

int mrmsmt_icls_create(const mrmsmt_shared* s, mricls_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!s || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrmsmt_icls_create: null argument");
    return make_ctx(s->shared->problem, out);
  });
}

}  // extern "C"
