// Device-side description of a batched ICLS problem (Math::ICLS::Solver, one solve per voxel).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

// Preconditioned problem of constrained_least_squares.h:91-152 plus the Gram matrix of the
// constraints (all FP64, row-major):
struct IclsDeviceProblem {
  const double* b2d;   // [nmeas][n]         H L^-T: y_u = b2d^T b
  const double* Bt;    // [n][ncon_pad]      transposed unit-row constraints (c_u = B y_u - t)
  const double* t;     // [ncon_pad]         scaled constraint vector (zeros when absent, and as padding)
  const double* G;     // [ncon][ncon_pad]   Gram matrix B B^T (rows of B_act B_act^T, and c = c_u + G l)
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
  uint16_t* iters;         // [nvox] or null (0xFFFF skipped, 0xFFFE active set beyond the largest capacity)
  uint16_t* nactive;       // [nvox] or null
  unsigned long long* stats;    // [8] or null (see mricls.h)
  const double* b64;       // when set: double measurements in place of b, same layout
  double* x64;             // when set: double solutions in place of x
};

// Scratch of one solve: a work counter per pass and two queues of voxel indices that carry the
// voxels whose active set outgrew a pass's capacity to the next pass (ping-pong).
struct IclsScratch {
  unsigned long long* counters;   // [kIclsMaxPasses] work counters, [kIclsMaxPasses] queue lengths
  uint32_t* queue[2];             // [capacity] each
  size_t capacity;                // queue entries (>= the number of valid voxels of the call)
};

// Active-set capacities of the passes, in slots (one active constraint per slot).  A solve starts
// at the smallest capacity above the parameter count (or a chosen one) and re-solves the voxels
// that need more slots at each larger capacity in turn.
constexpr int kIclsCapacities[] = {16, 32, 48, 64, 96, 128};
constexpr int kIclsNumCapacities = 6;
constexpr int kIclsMaxPasses = 3;
constexpr int kIclsMaxActive = 128;
constexpr int kIclsMaxParams = 128;
constexpr int kIclsMaxCon = 1024;

// The capacity the first pass uses when none is chosen: the smallest above n + num_eq.
int icls_default_capacity(int n, int num_eq);
bool icls_valid_capacity(int kcap);
// Words of 64 constraints a launch is compiled for (4, 5, 8 or 16): ncon_pad = 64 times this,
// the padding constraints zero rows of B and G (their values stay 0 and never enter the argmin).
int icls_constraint_words(int ncon);

hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsScratch& S,
                       int first_capacity, hipStream_t stream);

// out[v][r] = sum_j H[r][j] x[v][j] (FP64 sums) for the problem matrix H given transposed, Ht [n][nmeas]:
// the predicted signal of each voxel's solution.
hipError_t icls_predict(const double* Ht, int n, int nmeas, const float* x, int ld_x, size_t nvox, float* out,
                        int ld_out, hipStream_t stream);

}  // namespace mrcsd
