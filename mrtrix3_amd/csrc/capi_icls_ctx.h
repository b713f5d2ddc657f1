// The ICLS solver context behind the mricls_* C ABI (shared by its translation units).
#pragma once

#include <stdint.h>

#include "capi_common.h"
#include "kernels/icls_solve.h"

struct mricls_ctx {
  int device = 0;
  mrcsd::IclsDeviceProblem P{};
  mrcsd::DeviceBuffer<double> b2d, Bt, t, G, L;
  // scratch of a solve: work counters and queue lengths of the passes, and the two queues of
  // voxels re-solved with more active-set slots (grown to the largest voxel count seen)
  mrcsd::DeviceBuffer<unsigned long long> counters;
  mrcsd::DeviceBuffer<uint32_t> queue;
  size_t queue_cap = 0;
  int capacity = 0;   // active-set slots of the first pass
};
