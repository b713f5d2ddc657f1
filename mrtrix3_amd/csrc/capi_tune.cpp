// Launch-shape autotuners and captured solve steps (include/mrtune.h): the persistent CSD grid and
// the ICLS first-pass capacity chosen by timing candidates on the caller's voxels, and one solve
// step -- optional warm-start inference, then the CSD solve -- captured as a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrtune.h"
#include "capi_common.h"
#include "capi_csd_ctx.h"
#include "capi_icls_ctx.h"
#include "kernels/icls_solve.h"

namespace {

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  Events() {
    MRCSD_HIP(hipEventCreate(&a));
    MRCSD_HIP(hipEventCreate(&b));
  }
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// Median time of `reps` runs of `step` on `stream` after one untimed run; an error status stops.
template <class F>
int time_step(F&& step, hipStream_t stream, int reps, float* ms) {
  Events ev;
  int rc = step();
  if (rc) return rc;
  std::vector<float> t;
  for (int r = 0; r < reps; ++r) {
    MRCSD_HIP(hipEventRecord(ev.a, stream));
    if ((rc = step())) return rc;
    MRCSD_HIP(hipEventRecord(ev.b, stream));
    MRCSD_HIP(hipEventSynchronize(ev.b));
    float x = 0.0f;
    MRCSD_HIP(hipEventElapsedTime(&x, ev.a, ev.b));
    t.push_back(x);
  }
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return MRCSD_OK;
}

void clear(mrtune_report* r) {
  if (!r) return;
  *r = mrtune_report{};
}

void record(mrtune_report* r, int value, float ms) {
  if (!r || r->n_candidates >= MRTUNE_MAX_CANDIDATES) return;
  r->values[r->n_candidates] = value;
  r->ms[r->n_candidates] = ms;
  ++r->n_candidates;
}

}  // namespace

struct mrcsd_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~mrcsd_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

extern "C" {

int mrcsd_set_groups(mrcsd_ctx* ctx, int warm, int groups) {
  if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_set_groups: null context");
  if (groups < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_set_groups: negative workgroup count");
  ctx->groups[warm ? 1 : 0] = groups;
  return MRCSD_OK;
}

int mrcsd_get_groups(const mrcsd_ctx* ctx, int warm) { return ctx ? ctx->groups[warm ? 1 : 0] : -1; }

int mrcsd_autotune(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
                   const mrcsd_warm* warm, float* fod, void* stream, mrtune_report* report) {
  return mrcsd::guarded([&]() -> int {
    clear(report);
    if (!ctx || !dw || !fod || nvox == 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_autotune: null or empty argument");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const bool w = warm && (warm->F0 || warm->neg);
    int full = 0;
    MRCSD_HIP(mrcsd::csd_full_grid(ctx->P.n_sh, w, &full));
    // the full grid and fractions of it: fewer resident wavefronts per SIMD trade latency hiding for
    // cache footprint; more than the full grid only queues workgroups behind the persistent ones
    std::vector<int> cand;
    for (int num : {8, 7, 6, 5, 4}) cand.push_back(std::max(1, full * num / 8));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int saved = ctx->groups[w ? 1 : 0];
    int best = full;
    float best_ms = 0.0f;
    for (int g : cand) {
      ctx->groups[w ? 1 : 0] = g;
      float ms = 0.0f;
      const int rc = time_step([&]() { return mrcsd_run(ctx, dw, ld_dw, valid, nvox, fod, nullptr, warm, stream); }, s,
                               3, &ms);
      if (rc) {
        ctx->groups[w ? 1 : 0] = saved;
        return rc;
      }
      record(report, g, ms);
      if (best_ms == 0.0f || ms < best_ms) {
        best_ms = ms;
        best = g;
      }
    }
    ctx->groups[w ? 1 : 0] = best == full ? 0 : best;
    if (report) {
      report->chosen = best;
      report->full = full;
    }
    return MRCSD_OK;
  });
}

int mricls_autotune(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox, float* x,
                    void* stream, mrtune_report* report) {
  return mrcsd::guarded([&]() -> int {
    clear(report);
    if (!ctx || !b || !x || nvox == 0) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_autotune: null or empty argument");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int smallest = mrcsd::icls_default_capacity(ctx->P.n, ctx->P.num_eq);
    const int saved = ctx->capacity;
    int best = smallest;
    float best_ms = 0.0f;
    for (int i = 0; i < mrcsd::kIclsNumCapacities; ++i) {
      const int cap = mrcsd::kIclsCapacities[i];
      if (cap < smallest) continue;
      ctx->capacity = cap;
      float ms = 0.0f;
      const int rc = time_step(
          [&]() { return mricls_solve(ctx, b, ld_b, valid, nvox, x, nullptr, nullptr, nullptr, stream); }, s, 3, &ms);
      if (rc) {
        ctx->capacity = saved;
        return rc;
      }
      record(report, cap, ms);
      if (best_ms == 0.0f || ms < best_ms) {
        best_ms = ms;
        best = cap;
      }
    }
    ctx->capacity = best;
    if (report) {
      report->chosen = best;
      report->full = smallest;
    }
    return MRCSD_OK;
  });
}

int mrcsd_graph_create(mrcsd_ctx* ctx, mrwarm_net* net, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
                       const int* dwis_dev, const int* b0_dev, int n_b0, uint64_t* negmask, int budget, float* fod,
                       const mrcsd_outputs* out, mrcsd_graph** graph) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx || !dw || !fod || !graph || nvox == 0)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_graph_create: null or empty argument");
    if (net && (!dwis_dev || !negmask))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_graph_create: a warm-start step needs dwis_dev and negmask");
    *graph = nullptr;
    MRCSD_HIP(hipSetDevice(ctx->device));
    hipStream_t s = nullptr;
    MRCSD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<void, void (*)(void*)> hold(s, [](void* p) { (void)hipStreamDestroy(static_cast<hipStream_t>(p)); });
    mrcsd_warm w{nullptr, negmask, budget};
    // one eager step first: any one-off preparation (the inference's weight layout) happens here,
    // outside the capture, and the arguments are checked
    if (net) {
      int rc = mrwarm_infer(net, dw, ld_dw, nvox, dwis_dev, b0_dev, n_b0, nullptr, negmask, s);
      if (rc) return rc;
    }
    int rc = mrcsd_run(ctx, dw, ld_dw, valid, nvox, fod, out, net ? &w : nullptr, s);
    if (rc) return rc;
    MRCSD_HIP(hipStreamSynchronize(s));
    auto g = std::make_unique<mrcsd_graph>();
    g->device = ctx->device;
    MRCSD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    if (net) rc = mrwarm_infer(net, dw, ld_dw, nvox, dwis_dev, b0_dev, n_b0, nullptr, negmask, s);
    if (!rc) rc = mrcsd_run(ctx, dw, ld_dw, valid, nvox, fod, out, net ? &w : nullptr, s);
    hipGraph_t captured = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &captured);
    if (rc) {
      if (captured) (void)hipGraphDestroy(captured);
      return rc;
    }
    MRCSD_HIP(e);
    g->graph = captured;
    MRCSD_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *graph = g.release();
    return MRCSD_OK;
  });
}

int mrcsd_graph_launch(mrcsd_graph* graph, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!graph) return mrcsd::fail(MRCSD_ERR_ARG, "mrcsd_graph_launch: null graph");
    MRCSD_HIP(hipSetDevice(graph->device));
    MRCSD_HIP(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}

void mrcsd_graph_destroy(mrcsd_graph* graph) { delete graph; }

}  // extern "C"
