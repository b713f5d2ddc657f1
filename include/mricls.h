/*
 * mricls — batched inequality-constrained least squares and multi-tissue CSD set-up on the GPU
 * (part of libmrcsd.so; conventions as in mrcsd.h: row-major doubles, status codes,
 * mrcsd_last_error() for the message).
 *
 * Replaces, for a batch of right-hand sides:
 *   - Math::ICLS::Problem / Solver          (core/math/constrained_least_squares.h:73-365)
 *   - MSMT_CSD::Shared and the per-voxel MSMT_CSD solve
 *                                           (src/dwi/sdeconv/msmt_csd.h:47-275)
 *   - dwi2fod's MSMT_Processor              (cmd/dwi2fod.cpp:181-231)
 */
#ifndef MRICLS_H
#define MRICLS_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mricls_ctx mricls_ctx;

/* Problem(H, A, t, num_eq, sol_reg, con_reg, max_iter, tol, standard_form)
 * (constrained_least_squares.h:91-152).  H [nmeas][n] (or [n][n] with H^T H in the lower
 * triangle when standard_form), A [ncon][n] with the num_eq equality rows last, t [ncon] or NULL
 * (zeros).  Limits: n <= 128, ncon <= 1024, at most 128 simultaneously active constraints (the
 * reference has none; a voxel exceeding 128 reports iters 0xFFFE and mricls_solve_host fails). */
int mricls_create(const double* H, int nmeas, int n, const double* A, int ncon, const double* t,
                  int num_eq, double sol_reg, double con_reg, int max_iter, double tol,
                  int standard_form, mricls_ctx** out);
void mricls_destroy(mricls_ctx* ctx);
/* Dimensions and the effective max_niter (10 n when 0 was given, :105). */
int mricls_dims(const mricls_ctx* ctx, int* nmeas, int* n, int* ncon, int* max_niter);

/* Active-set slots of the first pass of mricls_solve (no counterpart in the reference, whose
 * active set grows without bound): a voxel whose active set outgrows them is solved again, from
 * the start and with the same arithmetic, with 64 and then 128 slots.  slots: 16, 32, 48, 64, 96
 * or 128; 0 restores the default, the smallest above n + num_eq.  Fewer slots mean less on-chip
 * memory per voxel and more voxels in flight. */
int mricls_set_capacity(mricls_ctx* ctx, int slots);
int mricls_get_capacity(const mricls_ctx* ctx);

/* Solver::operator()(x, b) (:237-357) for nvox right-hand sides.  Device pointers:
 *   b       [nvox][ld_b] float, the first nmeas entries of each row
 *   valid   [nvox] uint8 or NULL: 0 -> x = 0 and iters = 0xFFFF
 *   x       [nvox][n] float
 *   iters   [nvox] uint16 or NULL: the returned niter (0xFFFE: more than 128 active constraints)
 *   nactive [nvox] uint16 or NULL: size of the final active set
 *   stats   uint64[8] or NULL, accumulated: {voxels, outer iterations, inner solves,
 *           sum k, sum k^2, sum k^3 over inner solves, sum k over constraint updates,
 *           voxels over 128 active constraints}
 * Asynchronous on `stream`.  The first call with more voxels than any before grows the context's
 * voxel queues (a device allocation, after synchronising `stream`). */
int mricls_solve(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                 float* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream);
/* mricls_solve without any device allocation (no counterpart in the reference: for callers that
 * cannot allocate -- code under stream capture, serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as mricls_solve; its scratch (pass counters and the voxel
 * queues between capacities, 8 MiB per device, held until the process ends) is a fixed device
 * array shared by every call on the device: calls that overlap on one device race, so serialise
 * them.  At most MRICLS_NOALLOC_MAX_VOXELS voxels per call (MRCSD_ERR_ARG beyond). */
#define MRICLS_NOALLOC_MAX_VOXELS (1u << 20)
int mricls_solve_noalloc(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                         float* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream);

/* Host pointers (copies in, solves, copies out; synchronous).  Fails with MRCSD_ERR_SETUP when a
 * voxel needs more than 128 active constraints (its outputs are still written). */
int mricls_solve_host(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                      float* x, uint16_t* iters, uint16_t* nactive);

/* The predicted signal of each voxel's solution, out[v] = H x[v] for the problem matrix H
 * (MSMT_Processor's -predicted_signal, cmd/dwi2fod.cpp:216-220), FP64 sums rounded to float.
 * Device pointers x [nvox][ld_x] (the first n values of a row) and out [nvox][ld_out]; asynchronous
 * on `stream`.  Not available for a problem given in standard form (MRCSD_ERR_ARG). */
int mricls_predict(mricls_ctx* ctx, const float* x, int ld_x, size_t nvox, float* out, int ld_out, void* stream);

/* mricls_solve and mricls_solve_host with double-precision measurements and solutions, as
 * Solver<double>::operator()(x, b) takes and returns them (:237-357; the solve itself is FP64 in
 * both forms).  For right-hand sides that are not single-precision data, e.g. amp2response's
 * accumulated normal equations (cmd/amp2response.cpp:357-361). */
int mricls_solve_f64(mricls_ctx* ctx, const double* b, int ld_b, const uint8_t* valid, size_t nvox,
                     double* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream);
int mricls_solve_host_f64(mricls_ctx* ctx, const double* b, int ld_b, const uint8_t* valid, size_t nvox,
                          double* x, uint16_t* iters, uint16_t* nactive);

typedef struct mrmsmt_shared mrmsmt_shared;

/* MSMT_CSD::Shared(header) + parse_cmdline_options() + set_responses() + init().
 * grad [nvol][4] raw table; responses: tissue t is a [resp_rows[t]][resp_cols[t]] row-major block,
 * blocks concatenated; lmax [n_lmax] (the -lmax list) or NULL = from the responses capped at 8
 * (msmt_csd.h:102-114; n_lmax != ntissue is the reference's error);
 * directions NULL = built-in 300; norm_lambda / neg_lambda < 0 = defaults (1e-10);
 * shells [n_shells] the b-values of a -shells option or NULL = every shell (msmt_csd.h:49-54 with
 * DWI::ShellsOption, shells.cpp:100-281): volumes outside the selected shells get zero rows. */
int mrmsmt_shared_create(const double* grad, int nvol, const double* responses,
                         const int* resp_rows, const int* resp_cols, int ntissue, const int* lmax,
                         int n_lmax, const double* directions, int n_directions,
                         double norm_lambda, double neg_lambda, const double* shells, int n_shells,
                         mrmsmt_shared** out);
void mrmsmt_shared_destroy(mrmsmt_shared* s);
/* nshell, nparams (sum of NforL(lmax_t)), ncon (constraint rows), lmax_out[ntissue]. */
int mrmsmt_shared_info(const mrmsmt_shared* s, int* nshell, int* nparams, int* ncon, int* lmax_out);
/* which: 0 C [nvol][nparams], 1 A [ncon][nparams] */
int mrmsmt_shared_matrix(const mrmsmt_shared* s, int which, double* out, size_t capacity);
/* The GPU solver of its ICLS problem. */
int mrmsmt_icls_create(const mrmsmt_shared* s, mricls_ctx** out);

#ifdef __cplusplus
}
#endif
#endif
