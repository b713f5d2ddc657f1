// Device-side description of a batched ICLS problem (Math::ICLS::Solver, one solve per voxel).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

// Preconditioned problem of constrained_least_squares.h:91-152 plus derived matrices that let
// the solver skip per-voxel triangular solves (all FP64, row-major):
struct IclsDeviceProblem {
  const double* b2d;   // [nmeas][n]         H L^-T: y_u = b2d^T b
  const double* Bt;    // [n][ncon_pad]      transposed unit-row constraints (c_u = B y_u - t)
  const double* t;     // [ncon]             scaled constraint vector (zeros when absent)
  const double* G;     // [ncon][ncon_pad]   Gram matrix B B^T (B_act B_act^T, and c = c_u + G l)
  const double* L;     // [n][n]             lower Cholesky factor of H^T H (+ regularisation): x = L^-T y
  double lambda_min;   // constraint min-norm regularisation added to B_act B_act^T
  double tol;
  int nmeas, n, ncon, ncon_pad, num_eq, max_niter;
};

struct IclsLaunch {
  const float* b;          // [nvox][ld_b]: the first nmeas values of each row are the measurements
  int ld_b;
  const uint8_t* valid;    // [nvox] or null
  size_t nvox;
  float* x;                // [nvox][n]
  uint16_t* iters;         // [nvox] or null (0xFFFF skipped, 0xFFFE active set over capacity)
  uint16_t* nactive;       // [nvox] or null
  unsigned long long* stats;    // [4] or null: voxels, outer iterations, inner solves, sum k^3
  unsigned long long* counter;  // work counter
};

constexpr int kIclsMaxActive = 64;
constexpr int kIclsMaxParams = 64;
constexpr int kIclsMaxMeas = 512;
constexpr int kIclsMaxCon = 512;

hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, hipStream_t stream);

}  // namespace mrcsd
