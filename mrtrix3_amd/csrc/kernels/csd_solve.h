// Device-side problem description and launch parameters of the batched CSD solver.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

// All matrices FP64 in device memory (see csd_engine.cpp for how they are laid out).
struct CsdDeviceProblem {
  const double* MtM;     // [n_sh][n_sh] symmetric, norm lambda on the diagonal
  const double* HR;      // [n_hr][n_sh] constraint rows (row gathers for the normal matrix)
  const double* HRt;     // [n_hr_pad / NT][n_sh][NT]: direction r = b * NT + t of block b at (b, j, t),
                         // zero padded (NT = csd_threads_for(n_sh)); coalesced amplitude GEMV
  const double* M;       // [n_dw][n_sh] forward convolution (Mt_b = M^T b)
  const double* rconvT;  // [n_dw][n_init] initialisation matrix, transposed
  const int* dwis;       // [n_dw] volume index of each selected DW signal
  double thr;            // scaled threshold
  int n_dw, n_sh, n_init, n_hr, n_hr_pad, niter;
};

struct CsdLaunch {
  const float* dw;       // [nvox][ld_dw]
  int ld_dw;
  const uint8_t* valid;  // [nvox] or null
  size_t nvox;
  float* fod;            // [nvox][n_sh]
  uint16_t* iters;       // [nvox] or null
  uint16_t* nneg;        // [nvox] or null
  unsigned long long* stats;    // [6] or null
  const float* warm_F0;  // [nvox][n_sh] or null
  const unsigned long long* warm_neg;   // [nvox][n_hr_pad / 64] proposed negative sets, or null
  int warm_budget;       // iterate() calls allowed from the warm start before the plain restart
  unsigned long long* negmask;  // [nvox][n_hr_pad / 64] or null: neg set of the last iterate()
  unsigned long long* counter;  // work counter (device scalar)
  int max_groups;        // 0 = fill the device
};

hipError_t csd_launch(const CsdDeviceProblem& P, const CsdLaunch& L, hipStream_t stream);
// Workgroups of a full persistent grid (CUs x resident workgroups per CU) for the instantiation a
// launch with (warm) or without (!warm) a proposal runs.
hipError_t csd_full_grid(int n_sh, bool warm, int* groups);
bool csd_supported_nsh(int n_sh);
int csd_threads_for(int n_sh);

}  // namespace mrcsd
