/*
 * mrgt — global tractography's external energy on the GPU (part of libmrcsd.so; conventions as in
 * mrcsd.h).  Replaces, for a batch of voxels:
 *   - ExternalEnergyComputer's kernel matrices K and Ak and its NNLS problem
 *                                  (src/dwi/tractography/GT/externalenergy.cpp:30-112)
 *   - calcEnergy over the voxels, as resetEnergy runs it (:120-138, :245-252): per voxel
 *     y = dwi - K t, the NNLS fit f of y by Ak (ICLS::Problem(Ak, I), Solver<double>), the energy
 *     |y - Ak[:, 1:] f[1:]|^2 / nrows + mu t[0] (mu = ppot * sqrt(4 pi)), the isotropic fractions
 *     f[1:], and the sum of the energies.
 * The Metropolis-Hastings proposals of the tractography (add / eval / acceptChanges) change a few
 * voxels at a time on the host; mrgt_energy serves their batched re-evaluation and the full
 * resetEnergy pass.
 */
#ifndef MRGT_H
#define MRGT_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrgt_ctx mrgt_ctx;

/* grad [nvol][4] raw gradient table (through get_DW_scheme); resp_wm [wm_rows][wm_cols] zonal
 * white-matter response per shell; resp_iso [nf][nshell] isotropic responses (NULL when nf = 0);
 * lmax of the TOD (Properties::Lmax); weight (Properties::weight, scales K); ppot (the
 * Properties::ppot of mu).  Errors: "WM kernel size does not match the no. b-values in the image.",
 * "Isotropic kernel size does not match the no. b-values in the image." (externalenergy.cpp:62-67). */
int mrgt_create(const double* grad, int nvol, const double* resp_wm, int wm_rows, int wm_cols, const double* resp_iso,
                int nf, int lmax, double weight, double ppot, mrgt_ctx** out);
void mrgt_destroy(mrgt_ctx* ctx);
/* The host set-up alone (no GPU needed): K [nvol][NforL(lmax)] and Ak [nvol][nf + 1] into
 * caller buffers of those sizes. */
int mrgt_kernels(const double* grad, int nvol, const double* resp_wm, int wm_rows, int wm_cols, const double* resp_iso,
                 int nf, int lmax, double weight, double* K, double* Ak);
/* nrows (volumes), ncols (TOD coefficients), nf (isotropic tissues). */
int mrgt_dims(const mrgt_ctx* ctx, int* nrows, int* ncols, int* nf);
/* which: 0 K [nrows][ncols], 1 Ak [nrows][nf + 1] (row-major doubles). */
int mrgt_matrix(const mrgt_ctx* ctx, int which, double* out, size_t capacity);

/* calcEnergy for nvox voxels.  Device pointers: dwi [nvox][ld_dwi] float (the first nrows values
 * of a row), tod [nvox][ld_tod] float (the first ncols), eext [nvox] float (each voxel's energy, as
 * the eext image holds it), fiso [nvox][nf] float or NULL, total: one double (device), the sum of
 * the voxels' energies (in a fixed order).  Asynchronous on `stream`; the first call with more
 * voxels than any before grows the context's workspace (synchronising `stream`). */
int mrgt_energy(mrgt_ctx* ctx, const float* dwi, int ld_dwi, const float* tod, int ld_tod, size_t nvox, float* eext,
                float* fiso, double* total, void* stream);
/* The same with host pointers (synchronous). */
int mrgt_energy_host(mrgt_ctx* ctx, const float* dwi, int ld_dwi, const float* tod, int ld_tod, size_t nvox,
                     float* eext, float* fiso, double* total);

#ifdef __cplusplus
}
#endif
#endif
