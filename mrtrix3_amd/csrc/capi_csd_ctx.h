// The CSD solver context behind the mrcsd_* C ABI (shared by its translation units).
#pragma once

#include "capi_common.h"
#include "kernels/csd_solve.h"

struct mrcsd_ctx {
  int device = 0;
  mrcsd::CsdDeviceProblem P{};
  mrcsd::DeviceBuffer<double> MtM, HR, HRt, M, rconvT;
  mrcsd::DeviceBuffer<int> dwis;
  mrcsd::DeviceBuffer<unsigned long long> counter;
  int max_vol = 0;   // largest selected volume index (checked against ld_dw)
  // The warm-start certificate (neg(F(S)) == S implies the plain answer) holds when the fixed point
  // is the unique minimiser of the CSD objective: threshold 0 and a positive definite Mt_M.
  bool warm_certifiable = false;
  // persistent workgroups per launch chosen by mrcsd_autotune / mrcsd_set_groups (0: the full grid)
  int groups[2] = {0, 0};   // plain, warm-start instantiation
};
