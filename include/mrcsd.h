/*
 * mrcsd — C ABI of the MI355X constrained spherical deconvolution library (libmrcsd.so).
 *
 * Replaces the per-voxel solvers that MRtrix3's dwi2fod links in as header-only C++:
 *   - DWI::SDeconv::CSD::Shared / CSD::set / CSD::iterate   (src/dwi/sdeconv/csd.h:49-280)
 *   - Math::ICLS::Problem / Solver                          (core/math/constrained_least_squares.h:73-365)
 *   - the dwi2fod per-voxel processors                      (cmd/dwi2fod.cpp:120-231)
 * MRtrix3 has no plugin/FFI layer; a maintainer swaps the ThreadedLoop over CSD_Processor /
 * MSMT_Processor for one batched call per volume.
 *
 * Conventions: plain C types only; matrices are row-major double; every function returns
 * MRCSD_OK (0) or a negative error code and leaves a message for mrcsd_last_error() (thread
 * local).  No C++ exception crosses this boundary.  Device pointers are HIP device memory on the
 * current device; `stream` is a hipStream_t (NULL = default stream).  Contexts are not
 * thread-safe: use one per host thread, or serialise calls.
 */
#ifndef MRCSD_H
#define MRCSD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRCSD_OK 0
#define MRCSD_ERR_ARG -1      /* invalid argument / dimensions (message says which) */
#define MRCSD_ERR_HIP -2      /* a HIP runtime call failed */
#define MRCSD_ERR_SETUP -3    /* problem set-up failed (the reference would throw MR::Exception) */
#define MRCSD_ERR_NOGPU -4    /* no usable gfx950 device */

const char* mrcsd_last_error(void);
const char* mrcsd_version(void);

/* ------------------------------------------------------------------------------------------ */
/* Host-side problem set-up: CSD::Shared (csd.h:52-198)                                         */
/* ------------------------------------------------------------------------------------------ */

typedef struct mrcsd_options {
  int lmax;                  /* -lmax; 0 = from the response, capped at 8 (csd.h:127) */
  const double* filter;      /* -filter coefficients per even degree; NULL = [1 1 1] (csd.h:136) */
  int n_filter;
  const double* directions;  /* -directions [n][2] (az, el); NULL = built-in 300 (csd.h:54) */
  int n_directions;
  double neg_lambda;         /* -neg_lambda, default 1   (csd.h:32) */
  double norm_lambda;        /* -norm_lambda, default 1  (csd.h:33) */
  double threshold;          /* -threshold, default 0    (csd.h:34), range [-1, 10] */
  int niter;                 /* -niter, default 50       (csd.h:35), range [0, 1000] */
  int bvalue_scaling;        /* 0 auto, 1 on, 2 off      (gradient.cpp:280-324) */
  const double* shells;      /* -shells b-values; NULL = the largest shell (csd.h:57-60, shells.cpp:100-281) */
  int n_shells;
} mrcsd_options;

void mrcsd_options_default(mrcsd_options* opt);

typedef struct mrcsd_shared mrcsd_shared;

/* CSD::Shared(header) + parse_cmdline_options() + set_response() + init().
 * grad: [nvol][4] raw gradient table (x y z b); response: zonal coefficients [n_resp]. */
int mrcsd_shared_create(const double* grad, int nvol, const double* response, int n_resp,
                        const mrcsd_options* opt, mrcsd_shared** out);
void mrcsd_shared_destroy(mrcsd_shared* s);

typedef struct mrcsd_shared_info {
  int n_dw;          /* selected DW volumes (dwis.size()) */
  int n_sh;          /* nSH() */
  int n_init;        /* rows of rconv (NforL(lmax_response)) */
  int n_hr;          /* constraint directions */
  int lmax, lmax_response, lmax_data, niter;
  double threshold;  /* scaled by the constraint multiplier (csd.h:180) */
  double norm_lambda;/* value added to the diagonal of Mt_M (csd.h:192-195) */
} mrcsd_shared_info;

int mrcsd_shared_info_get(const mrcsd_shared* s, mrcsd_shared_info* info);
/* which: 0 rconv [n_init][n_dw], 1 M [n_dw][n_sh], 2 Mt_M [n_sh][n_sh] (full symmetric),
 *        3 HR_trans [n_hr][n_sh], 4 dwis (as doubles) [n_dw], 5 normalised grad [nvol][4] */
int mrcsd_shared_matrix(const mrcsd_shared* s, int which, double* out, size_t capacity);

/* ------------------------------------------------------------------------------------------ */
/* GPU CSD solver (CSD::set + CSD::iterate loop, dwi2fod.cpp:128-173, for a batch of voxels)    */
/* ------------------------------------------------------------------------------------------ */

typedef struct mrcsd_params {
  int n_dw, n_sh, n_init, n_hr;
  const double* rconv;     /* [n_init][n_dw] */
  const double* M;         /* [n_dw][n_sh] */
  const double* MtM;       /* [n_sh][n_sh] symmetric, norm lambda already on the diagonal */
  const double* HR_trans;  /* [n_hr][n_sh], constraint multiplier applied */
  const int* dwis;         /* [n_dw] volume index of each DW signal within an input row */
  double threshold;        /* scaled threshold */
  int niter;
} mrcsd_params;

typedef struct mrcsd_ctx mrcsd_ctx;

/* Copies the FP64 problem matrices to the current device. */
int mrcsd_create(const mrcsd_params* p, mrcsd_ctx** out);
int mrcsd_create_from_shared(const mrcsd_shared* s, mrcsd_ctx** out);
void mrcsd_destroy(mrcsd_ctx* ctx);

/* Optional warm start (see mrwarm.h; no counterpart in the reference).  Each voxel first runs
 * at most `budget` iterate() calls from the proposal: F0 replaces rconv*b (csd.h:238) and/or the
 * proposed negative set replaces the first amplitude evaluation.  A voxel whose negative set
 * reaches a fixed point there is accepted (for threshold 0 that is the unique minimiser, i.e. the
 * plain solver's answer); any other voxel is solved again from rconv*b with the full niter budget,
 * which reproduces the reference trajectory.  Ignored when niter is 0.
 * The certificate needs threshold 0 and a positive definite Mt_M (norm_lambda > 0, or M of full
 * column rank): mrcsd_run refuses a proposal otherwise (MRCSD_ERR_ARG; mrcsd_warm_supported tells).
 * One difference from the reference remains by design: a voxel the plain loop would leave
 * unconverged after niter iterations (csd.h:245-269 returns that iterate) is returned converged
 * when its proposal certifies. */
typedef struct mrcsd_warm {
  const float* F0;         /* device [nvox][n_sh] or NULL */
  const uint64_t* neg;     /* device [nvox][(n_hr+63)/64] proposed negative sets, or NULL */
  int budget;              /* iterate() calls allowed from the proposal (>= 2 to certify) */
} mrcsd_warm;

/* 1 when mrcsd_run accepts a warm-start proposal for this context (threshold 0 and Mt_M positive
 * definite), else 0. */
int mrcsd_warm_supported(const mrcsd_ctx* ctx);

/* Optional per-voxel outputs beyond the FOD. */
typedef struct mrcsd_outputs {
  uint16_t* iters;         /* [nvox] the dwi2fod loop counter n (n < niter <=> converged);
                              0xFFFF for skipped voxels */
  uint16_t* nneg;          /* [nvox] |neg| of the last iterate() */
  uint64_t* negmask;       /* [nvox][(n_hr+63)/64] bitmask of that neg set (bit r of word r/64) */
  uint64_t* stats;         /* uint64[6], accumulated: {voxels solved, solve iterations,
                              sum of |neg| over solves, converged voxels, warm-start fallbacks,
                              solves spent in warm-start passes} */
} mrcsd_outputs;

/* Solve nvox voxels.  Device pointers:
 *   dw     [nvox][ld_dw] float: all volumes of a voxel; the solver gathers dwis[], treats any
 *          non-finite selected value as "skip voxel" and clamps negatives to 0 (dwi2fod.cpp:156-173)
 *   valid  [nvox] uint8 or NULL: 0 = outside the mask -> zero output (dwi2fod.cpp:129-132)
 *   fod    [nvox][n_sh] float output
 *   out    NULL or optional outputs (each member nullable)
 *   warm   NULL or warm-start coefficients
 * Asynchronous on `stream`. */
int mrcsd_run(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
              float* fod, const mrcsd_outputs* out, const mrcsd_warm* warm, void* stream);

/* Same, with host pointers (copies in, solves, copies out; synchronous). */
int mrcsd_run_host(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
                   float* fod, uint16_t* iters, uint16_t* nneg);

#ifdef __cplusplus
}
#endif
#endif
