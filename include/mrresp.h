/* mrresp — amp2response on the GPU: the response function's normal equations accumulated over a
 * mask of single-fibre voxels (cmd/amp2response.cpp:152-191, Accumulator::operator()) and the
 * per-shell solve (cmd/amp2response.cpp:336-361): ordinary least squares, or the non-negativity
 * and monotonicity constrained problem on the batched ICLS kernel in standard form.
 * Error handling: every function returns MRCSD_OK or an error code; mrcsd_last_error() (mrcsd.h)
 * gives the message. */
#ifndef MRRESP_H
#define MRRESP_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrresp_ctx mrresp_ctx;

/* One shell: its unit gradient directions (host, [nvol][3], the Cartesian form of
 * DWI::gen_direction_matrix for the shell's volumes, amp2response.cpp:232-237 and :330), the
 * volume index of each (host, [nvol]) and the zonal harmonic degree (even, 0..12;
 * Accumulator::Shared, amp2response.cpp:116-128).  The directions must be finite for lmax > 0; at
 * lmax 0 they do not enter (a b=0 shell's zero directions give NaN ones, as the reference's). */
int mrresp_create(const double* dirs, const int* volumes, int nvol, int lmax, mrresp_ctx** out);
void mrresp_destroy(mrresp_ctx* ctx);
/* Number of zonal coefficients, lmax / 2 + 1. */
int mrresp_ncoef(const mrresp_ctx* ctx);

/* Accumulator::operator() over nvox voxels, summed as ~Accumulator does (amp2response.cpp:131-191).
 * Device pointers: amp [nvox][ld_amp] float (the DW signal, volume axis contiguous), fibre
 * [nvox][ld_fibre] float (the fibre direction in the first three values, normalised here), mask
 * [nvox] uint8 or NULL (every voxel).  Host outputs: M [n][n] (symmetric, n = mrresp_ncoef),
 * b [n], count (voxels in the mask).  Waits for the stream. */
int mrresp_accumulate(mrresp_ctx* ctx, const float* amp, int ld_amp, const float* fibre, int ld_fibre,
                      const uint8_t* mask, size_t nvox, double* M, double* b, uint64_t* count, void* stream);

/* The shell's response (amp2response.cpp:336-361): constrained = 0 or lmax = 0 solves
 * M rf = b (M.llt().solve(b)); otherwise the ICLS problem with M in standard form, amplitude and
 * elevation-derivative constraints at 0, 1, ..., 90 degrees (ZSH::init_amp_transform,
 * init_deriv_transform), regularisation 1e-10, solved by the GPU ICLS kernel (mricls.h).
 * Host arrays: M [n][n], b [n], rf [n] out; iterations (NULL allowed) gets the solver's count. */
int mrresp_solve(int lmax, const double* M, const double* b, int constrained, double* rf, int* iterations);

#ifdef __cplusplus
}
#endif

#endif
