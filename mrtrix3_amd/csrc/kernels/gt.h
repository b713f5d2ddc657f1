// Global tractography's external energy on the GPU (src/dwi/tractography/GT/externalenergy.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mrcsd {

// y[v][r] = dwi[v][r] - sum_c K[r][c] tod[v][c], FP64 (Kt = K transposed, [ncols][nrows]).
hipError_t gt_residual(const double* Kt, int nrows, int ncols, const float* dwi, int ld_dwi, const float* tod,
                       int ld_tod, size_t nvox, double* y, hipStream_t stream);
// e[v] = |y[v] - Ak[:, 1:] fk[v][1:]|^2 / nrows + mu tod[v][0] into eext (float) and the isotropic
// fractions fk[v][1:] into fiso (may be null); total = the sum of e over the voxels (FP64, fixed
// order: per workgroup, then over workgroups), written to *total (device).
hipError_t gt_energy(const double* Ak, int nrows, int nf, const double* y, const double* fk, const float* tod,
                     int ld_tod, double mu, size_t nvox, float* eext, float* fiso, double* partial,
                     size_t partial_cap, double* total, hipStream_t stream);
// Workgroups of gt_energy for nvox voxels (the size of its partial-sum buffer).
size_t gt_energy_blocks(size_t nvox);

}  // namespace mrcsd
