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

// The fused residual + NNLS + energy for n = nf + 1 <= 4 unknowns, when the voxel tiles fit the LDS
// ((nrows + ncols) x 64 floats <= 64 KB).  tab: for each subset s < 2^n of the unknowns, the n x n
// matrix holding (G_SS)^-1 on the subset's rows and columns and 0 elsewhere, then G = Ak^T Ak
// (n x n); Ak [nrows][n].  Same outputs as gt_residual + the NNLS + gt_energy.
bool gt_fused_fits(int nrows, int ncols, int n);
hipError_t gt_fused(const double* Kt, int nrows, int ncols, const double* Ak, int n, const double* tab,
                    const float* dwi, int ld_dwi, const float* tod, int ld_tod, double mu, size_t nvox, float* eext,
                    float* fiso, double* partial, size_t partial_cap, double* total, hipStream_t stream);
size_t gt_fused_blocks(size_t nvox);

}  // namespace mrcsd
