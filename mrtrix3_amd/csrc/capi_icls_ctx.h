// The ICLS solver context behind the mricls_* C ABI (shared by its translation units).
#pragma once

#include <stdint.h>

#include "capi_common.h"
#include "kernels/icls_solve.h"

namespace MR { namespace Math { namespace ICLS { class Problem; } } }

struct mricls_ctx {
  int device = 0;
  mrcsd::IclsDeviceProblem P{};
  mrcsd::DeviceBuffer<double> b2d, Bt, t, G, L;
  mrcsd::DeviceBuffer<double> Ht;   // [n][nmeas] problem matrix, transposed (predicted signals)
  bool standard_form = false;
  // scratch of a solve: work counters and queue lengths of the passes, and the two queues of
  // voxels re-solved with more active-set slots (grown to the largest voxel count seen)
  mrcsd::DeviceBuffer<unsigned long long> counters;
  mrcsd::DeviceBuffer<uint32_t> queue;
  size_t queue_cap = 0;
  int capacity = 0;   // active-set slots of the first pass
};

namespace mrcsd {
// A solver context for a problem set up on the host (mricls_create's second half); status code as
// the C ABI returns it.
int icls_ctx_from_problem(const MR::Math::ICLS::Problem& pr, mricls_ctx** out);
// mricls_solve / mricls_solve_f64 with the caller's scratch (per-thread callers) and input type.
int icls_solve_impl(mricls_ctx* ctx, const float* b, const double* b64, int ld_b, const uint8_t* valid,
                    size_t nvox, float* x, double* x64, uint16_t* iters, uint16_t* nactive, uint64_t* stats,
                    const IclsScratch* scratch, void* stream, const char* who);
}  // namespace mrcsd
