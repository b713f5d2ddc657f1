// C ABI of the learned warm start (include/mrwarm.h): network storage, forward / backward passes
// on the BF16 MFMA GEMM, Adam, and the feature / proposal kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/mrwarm.h"
#include "capi_common.h"
#include "host/npz.h"
#include "kernels/mlp.h"

namespace {

struct Layer {
  int fin = 0, fout = 0;
  mrcsd::DeviceBuffer<float> W, b, gW, gb, mW, vW, mb, vb;   // fp32 master, grads, Adam moments
  mrcsd::DeviceBuffer<__bf16> Wb;                          // bf16 copy used by the GEMMs
  mrcsd::DeviceBuffer<__bf16> WT;                          // [npad][kpad] transposed copy (fused inference)
  int kpad = 0, npad = 0;
};

}  // namespace

struct mrwarm_net {
  int n_in = 0, hidden = 0, nsh = 0, nlog = 0, step = 0;
  // the transposed copies WT of the fused inference hold the current weights (cleared whenever the
  // bf16 weights change: creation, set_params, a training step)
  bool wt_ready = false;
  Layer layers[4];             // n_in->h, h->h, h->h, h->(nsh+nlog)
  size_t cap = 0;              // batch capacity of the workspaces
  mrcsd::DeviceBuffer<float> pre[3], dact, out, dout, loss;
  mrcsd::DeviceBuffer<__bf16> act[3], dbf;
};

namespace {

int nout(const mrwarm_net* n) { return n->nsh + n->nlog; }

void ensure(mrwarm_net* n, size_t M) {
  if (M <= n->cap) return;
  const size_t h = size_t(n->hidden), o = size_t(nout(n));
  for (int l = 0; l < 3; ++l) {
    n->pre[l].alloc(M * h);
    n->act[l].alloc(M * h);
  }
  n->dact.alloc(M * std::max(h, o));
  n->out.alloc(M * o);
  n->dout.alloc(M * o);
  n->dbf.alloc(M * std::max(h, o));
  n->loss.alloc(2);
  n->cap = M;
}

mrcsd::GemmArgs gemm(const __bf16* A, long sAm, long sAk, const __bf16* B, long sBk, long sBn, float* C, long sCm,
                     long sCn, int M, int N, int K, int epi) {
  mrcsd::GemmArgs g{};
  g.A = A; g.sAm = sAm; g.sAk = sAk;
  g.B = B; g.sBk = sBk; g.sBn = sBn;
  g.C = C; g.sCm = sCm; g.sCn = sCn;
  g.M = M; g.N = N; g.K = K;
  g.epilogue = epi;
  return g;
}

void forward(mrwarm_net* n, const __bf16* x, size_t M, float* out, hipStream_t s) {
  const __bf16* in = x;
  for (int l = 0; l < 4; ++l) {
    Layer& L = n->layers[l];
    if (l < 3) {
      mrcsd::GemmArgs g = gemm(in, L.fin, 1, L.Wb.get(), L.fout, 1, n->pre[l].get(), L.fout, 1, int(M), L.fout,
                               L.fin, mrcsd::kEpiBiasGelu);
      g.bias = L.b.get();
      g.Cb = n->act[l].get();
      g.sCbm = L.fout;
      MRCSD_HIP(mrcsd::gemm_bf16(g, s));
      in = n->act[l].get();
    } else {
      mrcsd::GemmArgs g = gemm(in, L.fin, 1, L.Wb.get(), L.fout, 1, out, L.fout, 1, int(M), L.fout, L.fin,
                               mrcsd::kEpiBias);
      g.bias = L.b.get();
      MRCSD_HIP(mrcsd::gemm_bf16(g, s));
    }
  }
}

// Gradients of the warm-start loss into layers[*].gW / gb (fp32).
void backward(mrwarm_net* n, const __bf16* x, const float* tc, const uint8_t* tn, size_t M, float wc, float wl,
              hipStream_t s) {
  const int O = nout(n), H = n->hidden;
  forward(n, x, M, n->out.get(), s);
  MRCSD_HIP(hipMemsetAsync(n->loss.get(), 0, 2 * sizeof(float), s));
  MRCSD_HIP(mrcsd::warm_loss(n->out.get(), tc, tn, int(M), n->nsh, n->nlog, wc, wl, n->dout.get(), n->loss.get(), s));
  // top layer: dY = dout (fp32 -> bf16)
  MRCSD_HIP(mrcsd::to_bf16(n->dout.get(), n->dbf.get(), M * size_t(O), s));
  const float* dy_f = n->dout.get();
  for (int l = 3; l >= 0; --l) {
    Layer& L = n->layers[l];
    const __bf16* X = l ? n->act[l - 1].get() : x;
    const __bf16* dY = n->dbf.get();
    // dW = X^T dY   (A = X^T: element (k, m) at X[m * fin + k])
    MRCSD_HIP(mrcsd::gemm_bf16(gemm(X, 1, L.fin, dY, L.fout, 1, L.gW.get(), L.fout, 1, L.fin, L.fout, int(M),
                                    mrcsd::kEpiStore), s));
    MRCSD_HIP(mrcsd::colsum(dy_f, int(M), L.fout, L.gb.get(), s));
    if (l == 0) break;
    // dX = dY W^T into dact (fp32), then through GELU into dbf (bf16) and dact (fp32)
    MRCSD_HIP(mrcsd::gemm_bf16(gemm(dY, L.fout, 1, L.Wb.get(), 1, L.fout, n->dact.get(), L.fin, 1, int(M), L.fin,
                                    L.fout, mrcsd::kEpiStore), s));
    MRCSD_HIP(mrcsd::gelu_backward(n->pre[l - 1].get(), n->dact.get(), n->dbf.get(), n->dact.get(),
                                   M * size_t(H), s));
    dy_f = n->dact.get();
  }
}

}  // namespace

extern "C" {

int mrwarm_create(int n_in, int hidden, int nsh, int nlog, uint64_t seed, mrwarm_net** out) {
  return mrcsd::guarded([&]() -> int {
    if (n_in <= 0 || hidden <= 0 || nsh <= 0 || nlog < 0 || !out)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_create: bad dimensions");
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    auto n = std::make_unique<mrwarm_net>();
    n->n_in = n_in; n->hidden = hidden; n->nsh = nsh; n->nlog = nlog;
    const int dims[5] = {n_in, hidden, hidden, hidden, nsh + nlog};
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 4; ++l) {
      Layer& L = n->layers[l];
      L.fin = dims[l];
      L.fout = dims[l + 1];
      const size_t nw = size_t(L.fin) * L.fout;
      // Kaiming-uniform for the GELU layers, Xavier-uniform for the output layer
      const double bound = l < 3 ? std::sqrt(6.0 / L.fin) : std::sqrt(6.0 / (L.fin + L.fout));
      std::uniform_real_distribution<double> U(-bound, bound);
      std::vector<float> w(nw), b(size_t(L.fout), 0.0f);
      for (auto& v : w) v = float(U(rng));
      L.W.upload(w.data(), nw);
      L.b.upload(b.data(), b.size());
      L.Wb.alloc(nw);
      MRCSD_HIP(mrcsd::to_bf16(L.W.get(), L.Wb.get(), nw, nullptr));
      L.gW.alloc(nw); L.gb.alloc(b.size());
      std::vector<float> z(nw, 0.0f);
      L.mW.upload(z.data(), nw); L.vW.upload(z.data(), nw);
      L.mb.upload(z.data(), b.size()); L.vb.upload(z.data(), b.size());
    }
    MRCSD_HIP(hipDeviceSynchronize());
    *out = n.release();
    return MRCSD_OK;
  });
}

void mrwarm_destroy(mrwarm_net* net) { delete net; }

int mrwarm_dims(const mrwarm_net* net, int* n_in, int* hidden, int* nsh, int* nlog) {
  if (!net) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_dims: null network");
  if (n_in) *n_in = net->n_in;
  if (hidden) *hidden = net->hidden;
  if (nsh) *nsh = net->nsh;
  if (nlog) *nlog = net->nlog;
  return MRCSD_OK;
}

int mrwarm_file_dims(const char* path, int* n_in, int* hidden, int* nsh, int* nlog, size_t* n_params) {
  return mrcsd::guarded([&]() -> int {
    if (!path) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_file_dims: null path");
    const auto arrays = mrcsd::npz::load(path);
    const auto d = arrays.find("dims"), p = arrays.find("params");
    if (d == arrays.end() || p == arrays.end())
      return mrcsd::fail(MRCSD_ERR_ARG, std::string("\"") + path + "\" holds no 'dims' and 'params' arrays");
    const auto dims = d->second.as_double();
    if (dims.size() != 4 || p->second.descr != "<f4")
      return mrcsd::fail(MRCSD_ERR_ARG, std::string("\"") + path +
                                            "\": expected dims [n_in, hidden, nsh, nlog] and float32 params");
    if (n_in) *n_in = int(dims[0]);
    if (hidden) *hidden = int(dims[1]);
    if (nsh) *nsh = int(dims[2]);
    if (nlog) *nlog = int(dims[3]);
    if (n_params) *n_params = p->second.count();
    return MRCSD_OK;
  });
}

int mrwarm_load(const char* path, mrwarm_net** out) {
  return mrcsd::guarded([&]() -> int {
    if (!path || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_load: null argument");
    *out = nullptr;
    const auto arrays = mrcsd::npz::load(path);
    const auto d = arrays.find("dims"), p = arrays.find("params");
    if (d == arrays.end() || p == arrays.end())
      return mrcsd::fail(MRCSD_ERR_ARG, std::string("mrwarm_load: \"") + path + "\" holds no 'dims' and 'params' arrays");
    const auto dims = d->second.as_double();
    if (dims.size() != 4 || p->second.descr != "<f4")
      return mrcsd::fail(MRCSD_ERR_ARG, std::string("mrwarm_load: \"") + path +
                                            "\": expected dims [n_in, hidden, nsh, nlog] and float32 params");
    mrwarm_net* net = nullptr;
    int rc = mrwarm_create(int(dims[0]), int(dims[1]), int(dims[2]), int(dims[3]), 0, &net);
    if (rc) return rc;
    std::unique_ptr<mrwarm_net> hold(net);
    const size_t n = mrwarm_num_params(net);
    if (p->second.count() != n)
      return mrcsd::fail(MRCSD_ERR_ARG, std::string("mrwarm_load: \"") + path + "\" holds " +
                                            std::to_string(p->second.count()) + " parameters, the network has " +
                                            std::to_string(n));
    rc = mrwarm_set_params(net, reinterpret_cast<const float*>(p->second.data.data()), n);
    if (rc) return rc;
    *out = hold.release();
    return MRCSD_OK;
  });
}

size_t mrwarm_num_params(const mrwarm_net* net) {
  if (!net) return 0;
  size_t s = 0;
  for (const auto& L : net->layers) s += size_t(L.fin) * L.fout + size_t(L.fout);
  return s;
}

int mrwarm_get_params(const mrwarm_net* net, float* host, size_t n) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !host || n < mrwarm_num_params(net)) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_get_params: bad buffer");
    size_t o = 0;
    for (const auto& L : net->layers) {
      L.W.download(host + o, size_t(L.fin) * L.fout);
      o += size_t(L.fin) * L.fout;
      L.b.download(host + o, size_t(L.fout));
      o += size_t(L.fout);
    }
    return MRCSD_OK;
  });
}

int mrwarm_set_params(mrwarm_net* net, const float* host, size_t n) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !host || n < mrwarm_num_params(net)) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_set_params: bad buffer");
    size_t o = 0;
    for (auto& L : net->layers) {
      const size_t nw = size_t(L.fin) * L.fout;
      MRCSD_HIP(hipMemcpy(L.W.get(), host + o, nw * sizeof(float), hipMemcpyHostToDevice));
      o += nw;
      MRCSD_HIP(hipMemcpy(L.b.get(), host + o, size_t(L.fout) * sizeof(float), hipMemcpyHostToDevice));
      o += size_t(L.fout);
      MRCSD_HIP(mrcsd::to_bf16(L.W.get(), L.Wb.get(), nw, nullptr));
    }
    net->wt_ready = false;
    MRCSD_HIP(hipDeviceSynchronize());
    return MRCSD_OK;
  });
}

int mrwarm_features(const float* dw, int ld_dw, size_t M, const int* dwis_dev, int n_in, const int* b0_dev,
                    int n_b0, void* x_bf16, float* s0, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!dw || !dwis_dev || !x_bf16 || !s0 || n_in <= 0 || n_b0 < 0 || (n_b0 && !b0_dev))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_features: bad argument");
    MRCSD_HIP(mrcsd::features(dw, ld_dw, M, dwis_dev, n_in, b0_dev, n_b0, static_cast<__bf16*>(x_bf16), s0,
                              static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

int mrwarm_forward(mrwarm_net* net, const void* x_bf16, size_t M, float* out, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !x_bf16 || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_forward: null argument");
    if (M > size_t(1) << 31) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_forward: batch too large");
    ensure(net, M);
    forward(net, static_cast<const __bf16*>(x_bf16), M, out, static_cast<hipStream_t>(stream));
    return MRCSD_OK;
  });
}

int mrwarm_train_step(mrwarm_net* net, const void* x_bf16, const float* tgt_coef, const uint8_t* tgt_neg,
                      size_t M, float lr, float wc, float wl, float* loss_out, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !x_bf16 || !tgt_coef || !tgt_neg || M == 0)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_train_step: null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ensure(net, M);
    backward(net, static_cast<const __bf16*>(x_bf16), tgt_coef, tgt_neg, M, wc, wl, s);
    ++net->step;
    net->wt_ready = false;
    for (auto& L : net->layers) {
      MRCSD_HIP(mrcsd::adam_step(L.W.get(), L.Wb.get(), L.gW.get(), L.mW.get(), L.vW.get(), size_t(L.fin) * L.fout,
                                 lr, 0.9f, 0.999f, 1e-8f, net->step, 0.0f, s));
      MRCSD_HIP(mrcsd::adam_step(L.b.get(), nullptr, L.gb.get(), L.mb.get(), L.vb.get(), size_t(L.fout), lr, 0.9f,
                                 0.999f, 1e-8f, net->step, 0.0f, s));
    }
    if (loss_out) {
      MRCSD_HIP(hipMemcpyAsync(loss_out, net->loss.get(), 2 * sizeof(float), hipMemcpyDeviceToHost, s));
      MRCSD_HIP(hipStreamSynchronize(s));
    }
    return MRCSD_OK;
  });
}

int mrwarm_gradients(mrwarm_net* net, const void* x_bf16, const float* tgt_coef, const uint8_t* tgt_neg,
                     size_t M, float wc, float wl, float* grads_host, float* loss_out) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !x_bf16 || !tgt_coef || !tgt_neg || !grads_host || M == 0)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_gradients: null argument");
    ensure(net, M);
    backward(net, static_cast<const __bf16*>(x_bf16), tgt_coef, tgt_neg, M, wc, wl, nullptr);
    MRCSD_HIP(hipDeviceSynchronize());
    size_t o = 0;
    for (const auto& L : net->layers) {
      L.gW.download(grads_host + o, size_t(L.fin) * L.fout);
      o += size_t(L.fin) * L.fout;
      L.gb.download(grads_host + o, size_t(L.fout));
      o += size_t(L.fout);
    }
    if (loss_out) net->loss.download(loss_out, 2);
    return MRCSD_OK;
  });
}

int mrwarm_infer(mrwarm_net* net, const float* dw, int ld_dw, size_t M, const int* dwis_dev, const int* b0_dev,
                 int n_b0, float* F0, uint64_t* negmask, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!net || !dw || !dwis_dev || !negmask || (n_b0 > 0 && !b0_dev))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_infer: null argument");
    if (net->hidden != mrcsd::kInferHidden || net->n_in > mrcsd::kInferMaxIn || net->nsh + net->nlog > 512)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_infer: the fused path needs hidden = " +
                                                    std::to_string(mrcsd::kInferHidden) + ", n_in <= " +
                                                    std::to_string(mrcsd::kInferMaxIn) + ", nsh + nlog <= 512");
    if (ld_dw < 1) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_infer: ld_dw must be positive");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    mrcsd::InferArgs a{};
    for (int l = 0; l < 4; ++l) {
      Layer& L = net->layers[l];
      const int kpad = (L.fin + 31) / 32 * 32, npad = (L.fout + 15) / 16 * 16;
      if (L.kpad != kpad || L.npad != npad) {
        MRCSD_HIP(hipStreamSynchronize(s));   // an earlier launch may still read the old copy
        L.WT.alloc(size_t(kpad) * npad);
        L.kpad = kpad;
        L.npad = npad;
        net->wt_ready = false;
      }
      if (!net->wt_ready) MRCSD_HIP(mrcsd::transpose_pad(L.Wb.get(), L.fin, L.fout, L.WT.get(), kpad, npad, s));
      a.WT[l] = L.WT.get();
      a.ldw[l] = kpad;
      a.bias[l] = L.b.get();
    }
    net->wt_ready = true;
    a.dw = dw; a.ld = ld_dw; a.M = M;
    a.dwis = dwis_dev; a.n_in = net->n_in; a.k0 = net->layers[0].kpad;
    a.b0 = b0_dev; a.n_b0 = n_b0;
    a.nsh = net->nsh; a.nlog = net->nlog;
    a.F0 = F0;
    a.negmask = reinterpret_cast<unsigned long long*>(negmask);
    MRCSD_HIP(mrcsd::infer_propose(a, s));
    return MRCSD_OK;
  });
}

int mrwarm_propose(const float* out, const float* s0, size_t M, int nsh, int nlog, float* F0, uint64_t* negmask,
                   void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!out || !s0 || !F0 || !negmask) return mrcsd::fail(MRCSD_ERR_ARG, "mrwarm_propose: null argument");
    MRCSD_HIP(mrcsd::propose(out, s0, M, nsh, nlog, F0, reinterpret_cast<unsigned long long*>(negmask),
                             static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

}  // extern "C"
