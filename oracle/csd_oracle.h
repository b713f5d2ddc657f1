/*
 * CPU oracle for the constrained spherical deconvolution path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (MRtrix3 3.0.8), used by the tests as the
 * checker, by the benchmark as the CPU baseline ("port"), and by the smoke test.  Nothing in
 * the product links, loads or calls this code.  Every entry point cites the reference lines it
 * restates.  Arithmetic is carried in `orc_real` (double by default; the `long double` build,
 * -DORC_LONG_DOUBLE, bounds the FP64 rounding error of the double build).  All matrices crossing
 * this interface are row-major `double`.
 *
 * Pinning (how each function is checked independently of the reference binary, which cannot be
 * built here because Eigen 3 is absent):
 *   orc_plm_sph / orc_sh_transform : closed forms of Y_00, Y_2m, quadrature orthonormality,
 *                                    addition theorem                (tests/test_oracle_sh.py)
 *   orc_zsh2rh                     : closed form r_l * sqrt(4 pi / (2l+1))
 *   orc_pinv                       : Moore-Penrose identities, numpy.linalg.pinv
 *   orc_csd_init                   : structural closed forms (M = SHT diag(RH), c = 50 r0 / R,
 *                                    lambda = 2e-4 (M^T M)_00, +0.01 diagonal quirk)
 *   orc_csd_solve                  : stationarity of Phi(F) at convergence, brute-force convex
 *                                    minimisation on tiny problems, -niter 0 == rconv b
 *   orc_icls_*                     : the reference's Matlab quadprog golden vectors (1e-6),
 *                                    KKT conditions, brute-force active-set enumeration
 *   orc_shells_* / orc_dw_*        : rule-by-rule unit cases
 * Parity unpinned (needs the reference binary + its absent test data): exact trajectories and
 * final states of voxels that hit the iteration cap, and last-bit agreement with Eigen.
 */
#ifndef MRTRIX3_AMD_CSD_ORACLE_H
#define MRTRIX3_AMD_CSD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- spherical harmonics (core/math/legendre.h, SH.h, ZSH.h) ---- */
int    orc_nforl(int lmax);
int    orc_lforn(int n);
void   orc_plm_sph(double* array, int lmax, int m, double x);
void   orc_sh_transform(const double* dirs, int ndir, int lmax, double* sht);
void   orc_zsh2rh(const double* zsh, int n, double* rh);

/* ---- dense linear algebra (core/math/least_squares.h) ---- */
int    orc_pinv(const double* M, int rows, int cols, double* out);
int    orc_llt_solve(const double* A_lower, int n, const double* b, double* x);

/* ---- gradient table and shells (core/dwi/gradient.*, shells.*) ---- */
int    orc_dw_scheme_normalise(double* grad, int nvol, int scaling_mode);
void   orc_gen_direction_matrix(const double* grad, const int* vols, int nvol, double* dirs);
int    orc_shells_cluster(const double* bvals, int nvol, double bzero_thr, double eps, int* cluster);
int    orc_csd_select_volumes(const double* grad, int nvol, double bzero_thr, double eps,
                              int* vols_out, int* nvols_out);

/* ---- CSD::Shared (src/dwi/sdeconv/csd.h:49-212) ---- */
typedef struct orc_csd_shared {
  int n_dw, n_sh, n_init, n_hr, lmax, lmax_response, lmax_data, niter;
  double threshold;      /* already multiplied by the constraint multiplier */
  double norm_lambda;    /* the value added to the diagonal of Mt_M */
  double* rconv;         /* [n_init][n_dw] */
  double* M;             /* [n_dw][n_sh] */
  double* MtM;           /* [n_sh][n_sh], full symmetric copy, lambda on the diagonal */
  double* HR_trans;      /* [n_hr][n_sh] */
} orc_csd_shared;

int  orc_csd_init(const double* dw_dirs, int n_dw, const double* hr_dirs, int n_hr,
                  const double* response, int n_resp, int lmax_cmdline,
                  const double* init_filter, int n_filter,
                  double neg_lambda, double norm_lambda, double threshold, int niter,
                  orc_csd_shared** out, char* err, int errlen);
void orc_csd_free(orc_csd_shared* s);

/* One voxel through CSD::set + the dwi2fod iterate loop (dwi2fod.cpp:135-146).
 * b[n_dw] is the selected, clamped signal.  Writes F[n_sh]; returns the loop counter n
 * (n < niter <=> converged).  neg_out (nullable) gets the final neg list, *nneg its length.
 * trace_k (nullable, >= niter entries) gets |neg| of each solve. */
int  orc_csd_solve(const orc_csd_shared* s, const double* b, double* F,
                   int* neg_out, int* nneg, int* trace_k);

/* The same loop started from a given F0 (warm start); neg0 unused; niter from s. */
int  orc_csd_solve_from(const orc_csd_shared* s, const double* b, const double* F0, double* F,
                        int* neg_out, int* nneg, int* trace_k);

/* Threaded whole-volume CSD with the reference's scheduling: outer (y,z) lines dispensed under a
 * mutex to nthreads workers, each with its own workspace (core/algo/threaded_loop.h:324-364).
 * dw is [nz][ny][nx][nvol] float32 (volume axis contiguous), mask nullable [nz][ny][nx];
 * fod is [nz][ny][nx][n_sh] float32; iters nullable.  Returns 0. */
int  orc_csd_run_volume(const orc_csd_shared* s, const float* dw, int nvol, const int* vols,
                        const uint8_t* mask, int nx, int ny, int nz, int nthreads,
                        float* fod, int16_t* iters);

/* ---- Math::ICLS (core/math/constrained_least_squares.h:73-365) ---- */
typedef struct orc_icls_problem {
  int nmeas, n, ncon, num_eq, max_niter;
  double lambda_min_norm, tol;
  double* chol;   /* [n][n] lower factor L */
  double* b2d;    /* [nmeas][n] */
  double* B;      /* [ncon][n] */
  double* t;      /* [ncon] or NULL */
  double* H;      /* [nmeas][n] (copy, for predicted signal) */
} orc_icls_problem;

int  orc_icls_problem_create(const double* H, int nmeas, int n, const double* A, int ncon,
                             const double* t /*nullable*/, int num_eq, double sol_reg,
                             double con_reg, int max_iter, double tol, int standard_form,
                             orc_icls_problem** out, char* err, int errlen);
void orc_icls_problem_free(orc_icls_problem* p);
/* Returns niter; x[n]; active_out (nullable, [ncon]) gets the final active set. */
int  orc_icls_solve(const orc_icls_problem* p, const double* b, double* x, uint8_t* active_out);

/* Threaded batch of ICLS solves over voxels: b is [nvox][nmeas] float32, x [nvox][n] float32. */
int  orc_icls_run_batch(const orc_icls_problem* p, const float* b, size_t nvox,
                        const uint8_t* valid, int nthreads, float* x, int16_t* iters);

/* ---- MSMT_CSD::Shared::init (src/dwi/sdeconv/msmt_csd.h:100-218) ---- */
/* grad [nvol][4] (normalised), shell_of_vol[nvol] (shell index 0..nshell-1, sorted by b),
 * responses: tissue t has rows=nshell, cols=resp_cols[t] (row-major, concatenated),
 * lmax[t] per tissue.  Builds C [nvol][Σn_t] and A [Σm_t][Σn_t]. */
int  orc_msmt_build(const double* grad, int nvol, const int* shell_of_vol, int nshell,
                    const double* responses, const int* resp_cols, int ntissue, const int* lmax,
                    const double* hr_dirs, int n_hr, double* C, double* A, int* ncols, int* nrows_A);

const char* orc_real_name(void);

/* Single-thread timing of the rank-k normal-matrix update (k constraint rows, evenly spaced over the
 * directions) and the LLT of CSD::iterate, reps times each: total seconds of each into *t_syrk and
 * *t_llt.  FLOP: n(n+1)k for the update, n^3/3 for the factorisation. */
int orc_time_syrk_llt(const orc_csd_shared* s, int k, int reps, double* t_syrk, double* t_llt);

#ifdef __cplusplus
}
#endif
#endif
