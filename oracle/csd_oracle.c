/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see csd_oracle.h for scope and how each part is pinned).
 * Plain C restatement of MRtrix3 3.0.8's CSD / MSMT-CSD arithmetic, written for this project.
 * Nothing in the product links or calls this file.
 */
#define _GNU_SOURCE
#include "csd_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef ORC_LONG_DOUBLE
typedef long double orc_real;
#define RSQRT sqrtl
#define RCOS cosl
#define RSIN sinl
#define RACOS acosl
#define RATAN2 atan2l
#define RLOG logl
#define RFABS fabsl
#else
typedef double orc_real;
#define RSQRT sqrt
#define RCOS cos
#define RSIN sin
#define RACOS acos
#define RATAN2 atan2
#define RLOG log
#define RFABS fabs
#endif

const char* orc_real_name(void) { return sizeof(orc_real) == sizeof(double) ? "double" : "long double"; }

static void set_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) { strncpy(err, msg, (size_t)errlen - 1); err[errlen - 1] = 0; }
}

/* ------------------------------------------------------------------------------------------ */
/* SH index maps: core/math/SH.h:46-73, ZSH.h:42-57                                             */
/* ------------------------------------------------------------------------------------------ */
int orc_nforl(int lmax) { return (lmax + 1) * (lmax + 2) / 2; }
static int sh_index(int l, int m) { return l * (l + 1) / 2 + m; }
/* SH.h:70-73: evaluated with a *float* square root. */
int orc_lforn(int n) {
  if (!n) return 0;
  return 2 * (int)floorf((float)((sqrtf((float)(1 + 8 * n)) - 3.0) / 4.0));
}
static int zsh_nforl(int lmax) { return 1 + lmax / 2; }
static int zsh_lforn(int n) { return 2 * (n - 1); }

/* legendre.h:66-69: double-factorial ratio helper, recursion in steps of 2. */
static orc_real plm_helper(orc_real x, orc_real m) {
  return m < 1.0 ? 1.0 : x * (m - 1.0) / m * plm_helper(x, m - 2.0);
}

/* legendre.h:100-123: normalised associated Legendre functions for fixed m, l = m..lmax. */
static void plm_sph_r(orc_real* a, int lmax, int m, orc_real x) {
  orc_real x2 = x * x;
  if (m && x2 >= 1.0) {
    for (int n = m; n <= lmax; ++n) a[n] = 0.0;
    return;
  }
  a[m] = (orc_real)0.282094791773878;
  if (m) a[m] *= RSQRT((orc_real)(2 * m + 1) * plm_helper(1.0 - x2, 2.0 * m));
  if (m & 1) a[m] = -a[m];
  if (lmax == m) return;
  orc_real f = RSQRT((orc_real)(2 * m + 3));
  a[m + 1] = x * f * a[m];
  for (int n = m + 2; n <= lmax; n++) {
    a[n] = x * a[n - 1] - a[n - 2] / f;
    f = RSQRT((orc_real)(4 * n * n - 1) / (orc_real)(n * n - m * m));
    a[n] *= f;
  }
}

void orc_plm_sph(double* array, int lmax, int m, double x) {
  orc_real a[128];
  plm_sph_r(a, lmax, m, x);
  for (int n = m; n <= lmax; ++n) array[n] = (double)a[n];
}

/* SH.h:81-104: real even-order SH matrix; dirs are [az el] rows. */
static void sh_transform_r(const double* dirs, int ndir, int lmax, orc_real* sht) {
  const int nsh = orc_nforl(lmax);
  orc_real al[128];
  const orc_real sqrt2 = RSQRT((orc_real)2.0);
  for (int i = 0; i < ndir; i++) {
    orc_real z = RCOS((orc_real)dirs[2 * i + 1]);
    orc_real az = (orc_real)dirs[2 * i];
    orc_real* row = sht + (size_t)i * nsh;
    plm_sph_r(al, lmax, 0, z);
    for (int l = 0; l <= lmax; l += 2) row[sh_index(l, 0)] = al[l];
    for (int m = 1; m <= lmax; m++) {
      plm_sph_r(al, lmax, m, z);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) {
        row[sh_index(l, m)] = sqrt2 * al[l] * RCOS(m * az);
        row[sh_index(l, -m)] = sqrt2 * al[l] * RSIN(m * az);
      }
    }
  }
}

void orc_sh_transform(const double* dirs, int ndir, int lmax, double* sht) {
  const int nsh = orc_nforl(lmax);
  orc_real* t = (orc_real*)malloc(sizeof(orc_real) * (size_t)ndir * nsh);
  sh_transform_r(dirs, ndir, lmax, t);
  for (size_t i = 0; i < (size_t)ndir * nsh; i++) sht[i] = (double)t[i];
  free(t);
}

/* ZSH.h:220-231: rh_l = zsh_l / P_l^0(1). */
static void zsh2rh_r(const orc_real* zsh, int n, orc_real* rh) {
  const int lmax = zsh_lforn(n);
  orc_real al[128];
  plm_sph_r(al, lmax, 0, 1.0);
  for (int l = 0; l <= lmax; l += 2) rh[l / 2] = zsh[l / 2] / al[l];
}

void orc_zsh2rh(const double* zsh, int n, double* rh) {
  orc_real z[64], r[64];
  for (int i = 0; i < n; i++) z[i] = zsh[i];
  zsh2rh_r(z, n, r);
  for (int i = 0; i < n; i++) rh[i] = (double)r[i];
}

/* ------------------------------------------------------------------------------------------ */
/* Dense linear algebra                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* In-place Cholesky of the lower triangle of A (n x n, row-major); returns 0 on success,
 * k+1 if the k-th pivot is not positive (A is left partially factored, as Eigen's LLT does).
 * L_ij = (A_ij - sum_k L_ik L_jk) / L_jj with k ascending.  The strictly upper triangle is used as
 * scratch: row j past the diagonal receives column j of L, so the update of column j reads
 * earlier columns with unit stride and vectorises, with the same term order per entry. */
static int llt_inplace(orc_real* A, int n) {
  for (int j = 0; j < n; j++) {
    orc_real* rj = A + (size_t)j * n;
    orc_real d = rj[j];
    for (int k = 0; k < j; k++) d -= rj[k] * rj[k];
    if (!(d > 0)) return j + 1;
    d = RSQRT(d);
    rj[j] = d;
    for (int i = j + 1; i < n; i++) rj[i] = A[(size_t)i * n + j];
    for (int k = 0; k < j; k++) {
      const orc_real ljk = rj[k];
      const orc_real* ck = A + (size_t)k * n;   /* ck[i] = L_ik for i > k */
      for (int i = j + 1; i < n; i++) rj[i] -= ck[i] * ljk;
    }
    for (int i = j + 1; i < n; i++) {
      rj[i] /= d;
      A[(size_t)i * n + j] = rj[i];
    }
  }
  return 0;
}

/* Solve L L^T x = b in place with the factor in the lower triangle of L. */
static void llt_solve_inplace(const orc_real* L, int n, orc_real* x) {
  for (int i = 0; i < n; i++) {
    orc_real s = x[i];
    for (int k = 0; k < i; k++) s -= L[(size_t)i * n + k] * x[k];
    x[i] = s / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    orc_real s = x[i];
    for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * x[k];
    x[i] = s / L[(size_t)i * n + i];
  }
}

int orc_llt_solve(const double* A_lower, int n, const double* b, double* x) {
  orc_real* L = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * n);
  orc_real* y = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  for (size_t i = 0; i < (size_t)n * n; i++) L[i] = A_lower[i];
  for (int i = 0; i < n; i++) y[i] = b[i];
  int info = llt_inplace(L, n);
  llt_solve_inplace(L, n, y);
  for (int i = 0; i < n; i++) x[i] = (double)y[i];
  free(L); free(y);
  return info;
}

/* LDL^T with symmetric pivoting on the largest remaining diagonal entry (the strategy of
 * Eigen's LDLT, used by Math::pinv), then solve for nrhs right-hand sides in X [n][nrhs].
 * A is full symmetric storage and is overwritten: L (unit lower) is mirrored into the upper
 * triangle so that later symmetric row/column swaps permute it consistently; D on the diagonal. */
static void ldlt_solve(orc_real* A, int n, orc_real* X, int nrhs) {
  int* perm = (int*)malloc(sizeof(int) * (size_t)n);
  orc_real* v = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int k = 0; k < n; k++) {
    int p = k;
    orc_real best = RFABS(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (RFABS(A[(size_t)i * n + i]) > best) { best = RFABS(A[(size_t)i * n + i]); p = i; }
    if (p != k) {
      for (int j = 0; j < n; j++) { orc_real t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)p * n + j]; A[(size_t)p * n + j] = t; }
      for (int i = 0; i < n; i++) { orc_real t = A[(size_t)i * n + k]; A[(size_t)i * n + k] = A[(size_t)i * n + p]; A[(size_t)i * n + p] = t; }
      int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
    }
    const orc_real d = A[(size_t)k * n + k];
    for (int i = k + 1; i < n; i++) v[i] = A[(size_t)i * n + k];
    if (d != 0) {
      for (int i = k + 1; i < n; i++)
        for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= v[i] * v[j] / d;
    }
    for (int i = k + 1; i < n; i++) {
      orc_real lik = d != 0 ? v[i] / d : 0;
      A[(size_t)i * n + k] = lik;
      A[(size_t)k * n + i] = lik;
    }
  }
  orc_real* Y = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * nrhs);
  for (int i = 0; i < n; i++)
    for (int c = 0; c < nrhs; c++) Y[(size_t)i * nrhs + c] = X[(size_t)perm[i] * nrhs + c];
  for (int c = 0; c < nrhs; c++) {
    for (int i = 0; i < n; i++) {
      orc_real s = Y[(size_t)i * nrhs + c];
      for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * Y[(size_t)k * nrhs + c];
      Y[(size_t)i * nrhs + c] = s;
    }
    for (int i = 0; i < n; i++) {
      orc_real d = A[(size_t)i * n + i];
      Y[(size_t)i * nrhs + c] = (d != 0) ? Y[(size_t)i * nrhs + c] / d : 0;
    }
    for (int i = n - 1; i >= 0; i--) {
      orc_real s = Y[(size_t)i * nrhs + c];
      for (int k = i + 1; k < n; k++) s -= A[(size_t)k * n + i] * Y[(size_t)k * nrhs + c];
      Y[(size_t)i * nrhs + c] = s;
    }
  }
  for (int i = 0; i < n; i++)
    for (int c = 0; c < nrhs; c++) X[(size_t)perm[i] * nrhs + c] = Y[(size_t)i * nrhs + c];
  free(Y); free(perm); free(v);
}

/* least_squares.h:38-45: tall: (M^T M)^-1 M^T ; wide: ((M M^T)^-1 M)^T.  out is [cols][rows]. */
static void pinv_r(const orc_real* M, int rows, int cols, orc_real* out) {
  if (rows >= cols) {
    orc_real* G = (orc_real*)calloc((size_t)cols * cols, sizeof(orc_real));
    for (int i = 0; i < cols; i++)
      for (int j = 0; j < cols; j++) {
        orc_real s = 0;
        for (int r = 0; r < rows; r++) s += M[(size_t)r * cols + i] * M[(size_t)r * cols + j];
        G[(size_t)i * cols + j] = s;
      }
    for (int i = 0; i < cols; i++)
      for (int r = 0; r < rows; r++) out[(size_t)i * rows + r] = M[(size_t)r * cols + i];
    ldlt_solve(G, cols, out, rows);
    free(G);
  } else {
    orc_real* G = (orc_real*)calloc((size_t)rows * rows, sizeof(orc_real));
    orc_real* X = (orc_real*)malloc(sizeof(orc_real) * (size_t)rows * cols);
    for (int i = 0; i < rows; i++)
      for (int j = 0; j < rows; j++) {
        orc_real s = 0;
        for (int c = 0; c < cols; c++) s += M[(size_t)i * cols + c] * M[(size_t)j * cols + c];
        G[(size_t)i * rows + j] = s;
      }
    memcpy(X, M, sizeof(orc_real) * (size_t)rows * cols);
    ldlt_solve(G, rows, X, cols);
    for (int i = 0; i < rows; i++)
      for (int c = 0; c < cols; c++) out[(size_t)c * rows + i] = X[(size_t)i * cols + c];
    free(G); free(X);
  }
}

int orc_pinv(const double* M, int rows, int cols, double* out) {
  if (rows <= 0 || cols <= 0) return -1;
  orc_real* m = (orc_real*)malloc(sizeof(orc_real) * (size_t)rows * cols);
  orc_real* o = (orc_real*)malloc(sizeof(orc_real) * (size_t)rows * cols);
  for (size_t i = 0; i < (size_t)rows * cols; i++) m[i] = M[i];
  pinv_r(m, rows, cols, o);
  for (size_t i = 0; i < (size_t)rows * cols; i++) out[i] = (double)o[i];
  free(m); free(o);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Gradient table (core/dwi/gradient.h:77-94, gradient.cpp:280-341) and shells (shells.cpp)     */
/* ------------------------------------------------------------------------------------------ */

/* gradient.cpp:286-324.  scaling_mode: 0 auto, 1 force on, 2 force off.  Returns 1 if the
 * b-values were scaled by the squared norm, 0 otherwise. */
int orc_dw_scheme_normalise(double* grad, int nvol, int scaling_mode) {
  double maxlog = 0.0;
  double* sq = (double*)malloc(sizeof(double) * (size_t)(nvol > 0 ? nvol : 1));
  for (int r = 0; r < nvol; r++) {
    double* g = grad + 4 * (size_t)r;
    sq[r] = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (sq[r]) {
      double n = sqrt(sq[r]);
      g[0] /= n; g[1] /= n; g[2] /= n;
    }
    double v = sq[r] > 0.0 ? fabs(log(sq[r])) : 0.0;
    if (v > maxlog) maxlog = v;
  }
  int scale = (maxlog > 0.01 && scaling_mode == 0) || scaling_mode == 1;
  if (scale)
    for (int r = 0; r < nvol; r++) grad[4 * (size_t)r + 3] *= sq[r];
  free(sq);
  return scale;
}

/* gradient.h:77-94: [az el] of the selected rows. */
void orc_gen_direction_matrix(const double* grad, const int* vols, int nvol, double* dirs) {
  for (int i = 0; i < nvol; i++) {
    const double* g = grad + 4 * (size_t)vols[i];
    dirs[2 * i] = atan2(g[1], g[0]);
    double z = g[2] / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (z >= 1.0) dirs[2 * i + 1] = 0.0;
    else if (z <= -1.0) dirs[2 * i + 1] = M_PI;
    else dirs[2 * i + 1] = acos(z);
  }
}

/* shells.cpp:347-404 (clusterBvalues + regionQuery).  cluster[v] = 0 (unassigned), 1 (b=0 if
 * any), then 1.. / 2.. per shell in discovery order.  Returns the number of clusters. */
int orc_shells_cluster(const double* b, int nvol, double bzero_thr, double eps, int* cluster) {
  uint8_t* visited = (uint8_t*)calloc((size_t)nvol + 1, 1);
  int* nb = (int*)malloc(sizeof(int) * ((size_t)nvol * nvol + 1));
  int* nb2 = (int*)malloc(sizeof(int) * ((size_t)nvol + 1));
  int idx = 0;
  for (int i = 0; i < nvol; i++) cluster[i] = 0;
  for (int i = 0; i < nvol; i++)
    if (b[i] <= bzero_thr) { visited[i] = 1; idx = 1; cluster[i] = 1; }
  for (int i = 0; i < nvol; i++) {
    if (visited[i]) continue;
    visited[i] = 1;
    int nn = 0;
    for (int k = 0; k < nvol; k++)
      if (b[k] > bzero_thr && fabs(b[i] - b[k]) < eps) nb[nn++] = k;
    if (b[i] > bzero_thr && nn < 3) { cluster[i] = 0; continue; }
    cluster[i] = ++idx;
    for (int q = 0; q < nn; q++) {
      int j = nb[q];
      if (!visited[j]) {
        visited[j] = 1;
        int n2 = 0;
        for (int k = 0; k < nvol; k++)
          if (b[k] > bzero_thr && fabs(b[j] - b[k]) < eps) nb2[n2++] = k;
        if (n2 >= 3)
          for (int k = 0; k < n2; k++) nb[nn++] = nb2[k];
      }
      if (cluster[j] == 0) cluster[j] = idx;
    }
  }
  free(visited); free(nb); free(nb2);
  return idx;
}

/* Shells(grad).select_shells(true,false,true).largest().get_volumes() (csd.h:60), without a
 * -shells option: the largest-mean non-b=0 shell.  Returns 0, or -1 with no shell / -2 when the
 * shell count exceeds sqrt(nvol) (shells.cpp:306-307), -3 when only b=0 data exists. */
int orc_csd_select_volumes(const double* grad, int nvol, double bzero_thr, double eps,
                           int* vols_out, int* nvols_out) {
  double* b = (double*)malloc(sizeof(double) * (size_t)nvol);
  int* cl = (int*)malloc(sizeof(int) * (size_t)nvol);
  for (int i = 0; i < nvol; i++) b[i] = grad[4 * (size_t)i + 3];
  int nc = orc_shells_cluster(b, nvol, bzero_thr, eps, cl);
  if (nc < 1 || nc > sqrt((double)nvol)) { free(b); free(cl); return -2; }
  int best = -1; double best_mean = -1e300;
  int nonzero = 0;
  for (int c = 1; c <= nc; c++) {
    double s = 0; int cnt = 0;
    for (int i = 0; i < nvol; i++) if (cl[i] == c) { s += b[i]; cnt++; }
    if (!cnt) continue;
    double mean = s / cnt;
    if (mean <= bzero_thr) continue;
    nonzero++;
    if (mean > best_mean) { best_mean = mean; best = c; }
  }
  if (best < 0) { free(b); free(cl); return -3; }
  int n = 0;
  for (int i = 0; i < nvol; i++) if (cl[i] == best) vols_out[n++] = i;
  *nvols_out = n;
  free(b); free(cl);
  (void)nonzero;
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* CSD::Shared::init (csd.h:117-198)                                                            */
/* ------------------------------------------------------------------------------------------ */
int orc_csd_init(const double* dw_dirs, int n_dw, const double* hr_dirs, int n_hr,
                 const double* response, int n_resp, int lmax_cmdline,
                 const double* init_filter, int n_filter,
                 double neg_lambda, double norm_lambda, double threshold, int niter,
                 orc_csd_shared** out, char* err, int errlen) {
  const int lmax_data = orc_lforn(n_dw);                 /* csd.h:67 */
  int lmax_response = n_resp > 0 ? zsh_lforn(n_resp) : 0; /* csd.h:112 */
  if (lmax_data <= 0) { set_err(err, errlen, "data contain too few directions even for lmax = 2"); return -1; }
  if (lmax_response <= 0) { set_err(err, errlen, "response function does not contain anisotropic terms"); return -1; }
  int lmax = lmax_cmdline ? lmax_cmdline : (lmax_response < 8 ? lmax_response : 8);
  if (lmax <= 0 || lmax % 2) { set_err(err, errlen, "lmax must be a positive even integer"); return -1; }
  {
    int m = lmax_data < lmax ? lmax_data : lmax;
    if (m < lmax_response) lmax_response = m;
  }
  /* init_filter: default [1 1 1], resized (keep/zero-pad/truncate) to lmax_response/2+1 */
  const int nf = zsh_nforl(lmax_response);
  orc_real filt[64];
  for (int i = 0; i < nf; i++) {
    if (init_filter && n_filter > 0) filt[i] = i < n_filter ? init_filter[i] : 0.0;
    else filt[i] = i < 3 ? 1.0 : 0.0;
  }
  /* RH = ZSH2RH(response), zero-padded to lmax/2+1 */
  orc_real zr[64], rh[64];
  for (int i = 0; i < n_resp; i++) zr[i] = response[i];
  zsh2rh_r(zr, n_resp, rh);
  const int nrh_needed = zsh_nforl(lmax);
  for (int i = n_resp; i < nrh_needed; i++) rh[i] = 0.0;

  const int n_init = orc_nforl(lmax_response);
  const int n_sh = orc_nforl(lmax);
  orc_csd_shared* s = (orc_csd_shared*)calloc(1, sizeof(orc_csd_shared));
  s->n_dw = n_dw; s->n_sh = n_sh; s->n_init = n_init; s->n_hr = n_hr;
  s->lmax = lmax; s->lmax_response = lmax_response; s->lmax_data = lmax_data; s->niter = niter;
  s->rconv = (double*)malloc(sizeof(double) * (size_t)n_init * n_dw);
  s->M = (double*)malloc(sizeof(double) * (size_t)n_dw * n_sh);
  s->MtM = (double*)malloc(sizeof(double) * (size_t)n_sh * n_sh);
  s->HR_trans = (double*)malloc(sizeof(double) * (size_t)n_hr * n_sh);

  /* inverse sdeconv for initialisation (csd.h:145-158) */
  orc_real* fc = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_dw * n_init);
  sh_transform_r(dw_dirs, n_dw, lmax_response, fc);
  for (int i = 0; i < n_dw && i < n_init; i++) fc[(size_t)i * n_init + i] += 1.0e-2;
  orc_real* rc = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_init * n_dw);
  pinv_r(fc, n_dw, n_init, rc);
  {
    int l = 0, nl = 1;
    for (int row = 0; row < n_init; ++row) {
      if (row >= nl) { l++; nl = orc_nforl(2 * l); }
      orc_real f = filt[l] / rh[l];
      for (int c = 0; c < n_dw; c++) rc[(size_t)row * n_dw + c] *= f;
    }
  }
  for (size_t i = 0; i < (size_t)n_init * n_dw; i++) s->rconv[i] = (double)rc[i];
  free(fc); free(rc);

  /* forward convolution (csd.h:163-172) */
  orc_real* M = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_dw * n_sh);
  sh_transform_r(dw_dirs, n_dw, lmax, M);
  {
    int l = 0, nl = 1;
    for (int col = 0; col < n_sh; ++col) {
      if (col >= nl) { l++; nl = orc_nforl(2 * l); }
      for (int r = 0; r < n_dw; r++) M[(size_t)r * n_sh + col] *= rh[l];
    }
  }
  /* constraint directions (csd.h:175-180) */
  orc_real* H = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_hr * n_sh);
  sh_transform_r(hr_dirs, n_hr, lmax, H);
  orc_real cm = (orc_real)neg_lambda * 50.0 * (orc_real)response[0] / (orc_real)n_hr;
  for (size_t i = 0; i < (size_t)n_hr * n_sh; i++) s->HR_trans[i] = (double)(H[i] * cm);
  s->threshold = (double)((orc_real)threshold * cm);
  /* M and Mt_M (csd.h:184-195) */
  for (size_t i = 0; i < (size_t)n_dw * n_sh; i++) s->M[i] = (double)M[i];
  orc_real* G = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_sh * n_sh);
  for (int i = 0; i < n_sh; i++)
    for (int j = 0; j <= i; j++) {
      orc_real acc = 0;
      for (int r = 0; r < n_dw; r++) acc += M[(size_t)r * n_sh + i] * M[(size_t)r * n_sh + j];
      G[(size_t)i * n_sh + j] = G[(size_t)j * n_sh + i] = acc;
    }
  orc_real nl_ = 0;
  if (norm_lambda != 0.0) {
    nl_ = (orc_real)norm_lambda * (orc_real)0.0002 * G[0];
    for (int i = 0; i < n_sh; i++) G[(size_t)i * n_sh + i] += nl_;
  }
  s->norm_lambda = (double)nl_;
  for (size_t i = 0; i < (size_t)n_sh * n_sh; i++) s->MtM[i] = (double)G[i];
  free(M); free(H); free(G);
  *out = s;
  return 0;
}

void orc_csd_free(orc_csd_shared* s) {
  if (!s) return;
  free(s->rconv); free(s->M); free(s->MtM); free(s->HR_trans); free(s);
}

/* ------------------------------------------------------------------------------------------ */
/* CSD::set / CSD::iterate (csd.h:236-269) and the dwi2fod loop (dwi2fod.cpp:135-146)           */
/* ------------------------------------------------------------------------------------------ */
/* Per-thread workspace.  HRt (HR_trans transposed, [n_sh][n_hr]) and G (the gathered constraint
 * rows) let the amplitude pass and the rank-k update run as unit-stride loops the host compiler
 * vectorises; every sum keeps the reference's term order (see csd_iterate), so the results are
 * the same bit for bit as the plain dot-product loops. */
typedef struct csd_work {
  orc_real *F, *Mtb, *amps, *W, *Fd, *HRt, *G;
  int *neg, *old_neg;
  int nneg, nold;
} csd_work;

static void work_alloc(csd_work* w, const orc_csd_shared* s) {
  const int n = s->n_sh, R = s->n_hr;
  w->F = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  w->Mtb = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  w->amps = (orc_real*)malloc(sizeof(orc_real) * (size_t)R);
  w->W = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * n);
  w->Fd = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  w->HRt = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * R);
  w->G = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * R);
  w->neg = (int*)malloc(sizeof(int) * ((size_t)R + 1));
  w->old_neg = (int*)malloc(sizeof(int) * ((size_t)R + 1));
  for (int r = 0; r < R; r++)
    for (int j = 0; j < n; j++) w->HRt[(size_t)j * R + r] = (orc_real)s->HR_trans[(size_t)r * n + j];
}
static void work_free(csd_work* w) {
  free(w->F); free(w->Mtb); free(w->amps); free(w->W); free(w->Fd); free(w->HRt); free(w->G);
  free(w->neg); free(w->old_neg);
}

/* csd.h:236-243 (F0 nullable: the plain rconv initialisation). */
static void csd_set(const orc_csd_shared* s, csd_work* w, const double* b, const double* F0) {
  const int n = s->n_sh, N = s->n_dw;
  if (F0) {
    for (int i = 0; i < n; i++) w->F[i] = F0[i];
  } else {
    for (int i = 0; i < s->n_init; i++) {
      orc_real acc = 0;
      for (int c = 0; c < N; c++) acc += (orc_real)s->rconv[(size_t)i * N + c] * (orc_real)b[c];
      w->F[i] = acc;
    }
    for (int i = s->n_init; i < n; i++) w->F[i] = 0;
  }
  w->old_neg[0] = -1; w->nold = 1;
  for (int j = 0; j < n; j++) {
    orc_real acc = 0;
    for (int r = 0; r < N; r++) acc += (orc_real)s->M[(size_t)r * n + j] * (orc_real)b[r];
    w->Mtb[j] = acc;
  }
}

/* W = lower(Mt_M) + H_neg^T H_neg: W_ij = MtM_ij + sum_q h_qi h_qj with q ascending (csd.h:255-262),
 * accumulated one gathered row q at a time so the j loop is unit stride */
static void csd_form_normal(const orc_csd_shared* s, csd_work* w) {
  const int n = s->n_sh;
  const int k = w->nneg;
  for (int q = 0; q < k; q++) {
    const double* h = s->HR_trans + (size_t)w->neg[q] * n;
    orc_real* g = w->G + (size_t)q * n;
    for (int j = 0; j < n; j++) g[j] = (orc_real)h[j];
  }
  for (int i = 0; i < n; i++) {
    orc_real* wr = w->W + (size_t)i * n;
    for (int j = 0; j <= i; j++) wr[j] = (orc_real)s->MtM[(size_t)i * n + j];
    for (int q = 0; q < k; q++) {
      const orc_real* g = w->G + (size_t)q * n;
      const orc_real gi = g[i];
      for (int j = 0; j <= i; j++) wr[j] += gi * g[j];
    }
  }
}

/* csd.h:245-269.  Returns 1 on convergence (neg unchanged), else solves and returns 0. */
static int csd_iterate(const orc_csd_shared* s, csd_work* w) {
  const int n = s->n_sh, R = s->n_hr;
  /* a = HR_trans F: a_r = sum_j HR(r,j) F_j with j ascending, for all r at once */
  orc_real* const a = w->amps;
  for (int r = 0; r < R; r++) a[r] = 0;
  for (int j = 0; j < n; j++) {
    const orc_real fj = w->F[j];
    const orc_real* h = w->HRt + (size_t)j * R;
    for (int r = 0; r < R; r++) a[r] += h[r] * fj;
  }
  w->nneg = 0;
  for (int r = 0; r < R; r++)
    if (a[r] < (orc_real)s->threshold) w->neg[w->nneg++] = r;
  if (w->nneg == w->nold) {
    int same = 1;
    for (int i = 0; i < w->nneg; i++) if (w->neg[i] != w->old_neg[i]) { same = 0; break; }
    if (same) return 1;
  }
  csd_form_normal(s, w);
  llt_inplace(w->W, n);   /* the reference does not check LLT::info() (csd.h:264) */
  for (int i = 0; i < n; i++) w->F[i] = w->Mtb[i];
  llt_solve_inplace(w->W, n, w->F);
  memcpy(w->old_neg, w->neg, sizeof(int) * (size_t)w->nneg);
  w->nold = w->nneg;
  return 0;
}

static int csd_loop(const orc_csd_shared* s, csd_work* w, int* trace_k) {
  int n;
  for (n = 0; n < s->niter; n++) {
    if (csd_iterate(s, w)) break;
    if (trace_k) trace_k[n] = w->nneg;
  }
  return n;
}

static int csd_solve_common(const orc_csd_shared* s, const double* b, const double* F0, double* F,
                            int* neg_out, int* nneg, int* trace_k) {
  csd_work w;
  work_alloc(&w, s);
  csd_set(s, &w, b, F0);
  int n = csd_loop(s, &w, trace_k);
  for (int i = 0; i < s->n_sh; i++) F[i] = (double)w.F[i];
  if (neg_out || nneg) {
    /* final neg set of the returned F (equal to old_neg at convergence) */
    int k = 0;
    for (int r = 0; r < s->n_hr; r++) {
      orc_real a = 0;
      for (int j = 0; j < s->n_sh; j++) a += (orc_real)s->HR_trans[(size_t)r * s->n_sh + j] * w.F[j];
      if (a < (orc_real)s->threshold) { if (neg_out) neg_out[k] = r; k++; }
    }
    if (nneg) *nneg = k;
  }
  work_free(&w);
  return n;
}

int orc_csd_solve(const orc_csd_shared* s, const double* b, double* F, int* neg_out, int* nneg,
                  int* trace_k) {
  return csd_solve_common(s, b, NULL, F, neg_out, nneg, trace_k);
}

int orc_csd_solve_from(const orc_csd_shared* s, const double* b, const double* F0, double* F,
                       int* neg_out, int* nneg, int* trace_k) {
  return csd_solve_common(s, b, F0, F, neg_out, nneg, trace_k);
}

/* ---- threaded whole-volume runner (threaded_loop.h:324-364; dwi2fod.cpp:128-173) ---- */
typedef struct vol_job {
  const orc_csd_shared* s;
  const float* dw; int nvol; const int* vols; const uint8_t* mask;
  int nx, ny, nz;
  float* fod; int16_t* iters;
  pthread_mutex_t mtx;
  long next_line;
} vol_job;

static void* vol_worker(void* arg) {
  vol_job* J = (vol_job*)arg;
  const orc_csd_shared* s = J->s;
  csd_work w;
  work_alloc(&w, s);
  double* b = (double*)malloc(sizeof(double) * (size_t)s->n_dw);
  const long nlines = (long)J->ny * J->nz;
  for (;;) {
    pthread_mutex_lock(&J->mtx);
    long line = J->next_line++;
    pthread_mutex_unlock(&J->mtx);
    if (line >= nlines) break;
    for (int x = 0; x < J->nx; x++) {
      size_t v = (size_t)line * J->nx + x;
      float* out = J->fod + v * s->n_sh;
      int ok = !J->mask || J->mask[v];
      const float* src = J->dw + v * J->nvol;
      for (int i = 0; ok && i < s->n_dw; i++) {
        double d = src[J->vols[i]];
        if (!isfinite(d)) ok = 0;
        b[i] = d < 0.0 ? 0.0 : d;
      }
      if (!ok) {
        for (int i = 0; i < s->n_sh; i++) out[i] = 0.0f;
        if (J->iters) J->iters[v] = -1;
        continue;
      }
      csd_set(s, &w, b, NULL);
      int n = csd_loop(s, &w, NULL);
      for (int i = 0; i < s->n_sh; i++) out[i] = (float)w.F[i];
      if (J->iters) J->iters[v] = (int16_t)n;
    }
  }
  free(b);
  work_free(&w);
  return NULL;
}

int orc_csd_run_volume(const orc_csd_shared* s, const float* dw, int nvol, const int* vols,
                       const uint8_t* mask, int nx, int ny, int nz, int nthreads,
                       float* fod, int16_t* iters) {
  vol_job J;
  J.s = s; J.dw = dw; J.nvol = nvol; J.vols = vols; J.mask = mask;
  J.nx = nx; J.ny = ny; J.nz = nz; J.fod = fod; J.iters = iters; J.next_line = 0;
  pthread_mutex_init(&J.mtx, NULL);
  if (nthreads <= 0) {          /* -nthreads 0: run inline (threaded_loop.h:326-330) */
    vol_worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, vol_worker, &J);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&J.mtx);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Math::ICLS::Problem (constrained_least_squares.h:91-152)                                     */
/* ------------------------------------------------------------------------------------------ */
int orc_icls_problem_create(const double* H, int nmeas, int n, const double* A, int ncon,
                            const double* t, int num_eq, double sol_reg, double con_reg,
                            int max_iter, double tol, int standard_form,
                            orc_icls_problem** out, char* err, int errlen) {
  if (sol_reg < 0.0) { set_err(err, errlen, "FIXME: solution norm regularisation is negative (ICLS)"); return -1; }
  if (con_reg < 0.0) { set_err(err, errlen, "FIXME: constraint norm regularisation is negative (ICLS)"); return -1; }
  if (tol < 0.0) { set_err(err, errlen, "FIXME: tolerance is negative (ICLS)"); return -1; }
  orc_icls_problem* p = (orc_icls_problem*)calloc(1, sizeof(orc_icls_problem));
  p->nmeas = nmeas; p->n = n; p->ncon = ncon; p->num_eq = num_eq;
  p->max_niter = max_iter ? max_iter : 10 * n;
  p->lambda_min_norm = con_reg; p->tol = tol;
  orc_real* L = (orc_real*)calloc((size_t)n * n, sizeof(orc_real));
  if (standard_form) {
    for (int i = 0; i < n; i++) for (int j = 0; j <= i; j++) L[(size_t)i * n + j] = H[(size_t)i * n + j];
  } else {
    for (int i = 0; i < n; i++)
      for (int j = 0; j <= i; j++) {
        orc_real s = 0;
        for (int r = 0; r < nmeas; r++) s += (orc_real)H[(size_t)r * n + i] * (orc_real)H[(size_t)r * n + j];
        L[(size_t)i * n + j] = s;
      }
  }
  orc_real dmax = -INFINITY;
  for (int i = 0; i < n; i++) if (L[(size_t)i * n + i] > dmax) dmax = L[(size_t)i * n + i];
  for (int i = 0; i < n; i++) L[(size_t)i * n + i] += (orc_real)sol_reg * dmax;
  llt_inplace(L, n);
  for (int i = 0; i < n; i++) for (int j = i + 1; j < n; j++) L[(size_t)i * n + j] = 0;
  /* X L^T = Y  <=>  X = Y L^-T : solve row by row (forward substitution over columns) */
  const int nb2d = standard_form ? n : nmeas;
  orc_real* b2d = (orc_real*)malloc(sizeof(orc_real) * (size_t)nb2d * n);
  for (int r = 0; r < nb2d; r++) {
    for (int j = 0; j < n; j++) {
      orc_real y = standard_form ? (r == j ? 1.0 : 0.0) : (orc_real)H[(size_t)r * n + j];
      for (int k = 0; k < j; k++) y -= b2d[(size_t)r * n + k] * L[(size_t)j * n + k];
      b2d[(size_t)r * n + j] = y / L[(size_t)j * n + j];
    }
  }
  orc_real* B = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon * n);
  p->t = t ? (double*)malloc(sizeof(double) * (size_t)ncon) : NULL;
  for (int r = 0; r < ncon; r++) {
    for (int j = 0; j < n; j++) {
      orc_real y = (orc_real)A[(size_t)r * n + j];
      for (int k = 0; k < j; k++) y -= B[(size_t)r * n + k] * L[(size_t)j * n + k];
      B[(size_t)r * n + j] = y / L[(size_t)j * n + j];
    }
    orc_real nrm = 0;
    for (int j = 0; j < n; j++) nrm += B[(size_t)r * n + j] * B[(size_t)r * n + j];
    nrm = RSQRT(nrm);
    for (int j = 0; j < n; j++) B[(size_t)r * n + j] /= nrm;
    if (t) p->t[r] = (double)((orc_real)t[r] / nrm);
  }
  p->chol = (double*)malloc(sizeof(double) * (size_t)n * n);
  p->b2d = (double*)malloc(sizeof(double) * (size_t)nb2d * n);
  p->B = (double*)malloc(sizeof(double) * (size_t)ncon * n);
  p->H = (double*)malloc(sizeof(double) * (size_t)nb2d * n);
  for (size_t i = 0; i < (size_t)n * n; i++) p->chol[i] = (double)L[i];
  for (size_t i = 0; i < (size_t)nb2d * n; i++) p->b2d[i] = (double)b2d[i];
  for (size_t i = 0; i < (size_t)ncon * n; i++) p->B[i] = (double)B[i];
  if (standard_form) memcpy(p->H, H, sizeof(double) * (size_t)n * n);
  else memcpy(p->H, H, sizeof(double) * (size_t)nmeas * n);
  if (standard_form) p->nmeas = n;
  free(L); free(b2d); free(B);
  *out = p;
  return 0;
}

void orc_icls_problem_free(orc_icls_problem* p) {
  if (!p) return;
  free(p->chol); free(p->b2d); free(p->B); free(p->t); free(p->H); free(p);
}

/* ------------------------------------------------------------------------------------------ */
/* Math::ICLS::Solver::operator() (constrained_least_squares.h:237-357)                         */
/* ------------------------------------------------------------------------------------------ */
/* Per-thread workspace, allocated once per thread as the reference's Solver holds its vectors.
 * Bt (B transposed, [n][ncon]) and Gt (the active rows of B transposed, [n][na]) make every
 * product below a unit-stride loop the host compiler vectorises; each sum keeps the term order of
 * the plain dot-product form (ascending summation index), so results are the same bit for bit. */
typedef struct icls_ws {
  orc_real *y_u, *x, *c, *c_u, *lambda, *lambda_prev, *l, *BtB, *Bt, *Gt;
  int* act_idx;
  uint8_t* active;
} icls_ws;

static void icls_ws_alloc(icls_ws* w, const orc_icls_problem* P) {
  const int n = P->n, ncon = P->ncon;
  w->y_u = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  w->x = (orc_real*)malloc(sizeof(orc_real) * (size_t)n);
  w->c = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon);
  w->c_u = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon);
  w->lambda = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon);
  w->lambda_prev = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon);
  w->l = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon);
  w->BtB = (orc_real*)malloc(sizeof(orc_real) * (size_t)ncon * ncon);
  w->Bt = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * ncon);
  w->Gt = (orc_real*)malloc(sizeof(orc_real) * (size_t)n * ncon);
  w->act_idx = (int*)malloc(sizeof(int) * (size_t)ncon);
  w->active = (uint8_t*)malloc((size_t)ncon);
  for (int r = 0; r < ncon; r++)
    for (int j = 0; j < n; j++) w->Bt[(size_t)j * ncon + r] = (orc_real)P->B[(size_t)r * n + j];
}

static void icls_ws_free(icls_ws* w) {
  free(w->y_u); free(w->x); free(w->c); free(w->c_u); free(w->lambda); free(w->lambda_prev);
  free(w->l); free(w->BtB); free(w->Bt); free(w->Gt); free(w->act_idx); free(w->active);
}

/* c = B v - t: c_r = sum_j B(r,j) v_j (j ascending), then minus t_r */
static void icls_constraints(const orc_icls_problem* P, const icls_ws* w, const orc_real* v, orc_real* c) {
  const int n = P->n, ncon = P->ncon;
  for (int r = 0; r < ncon; r++) c[r] = 0;
  for (int j = 0; j < n; j++) {
    const orc_real vj = v[j];
    const orc_real* bt = w->Bt + (size_t)j * ncon;
    for (int r = 0; r < ncon; r++) c[r] += bt[r] * vj;
  }
  if (P->t)
    for (int r = 0; r < ncon; r++) c[r] -= (orc_real)P->t[r];
}

static int icls_solve_ws(const orc_icls_problem* P, icls_ws* w, const double* b, double* xout,
                         uint8_t* active_out) {
  const int n = P->n, ncon = P->ncon, nmeas = P->nmeas;
  const int num_eq = P->num_eq, num_ineq = ncon - num_eq;
  orc_real *y_u = w->y_u, *x = w->x, *c = w->c, *c_u = w->c_u, *lambda = w->lambda;
  orc_real *lambda_prev = w->lambda_prev, *l = w->l, *BtB = w->BtB, *Gt = w->Gt;
  int* act_idx = w->act_idx;
  uint8_t* active = w->active;
  for (int r = 0; r < ncon; r++) { lambda[r] = 0; lambda_prev[r] = 0; active[r] = 0; }

  /* y_u = b2d^T b: y_j = sum_r b2d(r,j) b_r (r ascending) */
  for (int j = 0; j < n; j++) y_u[j] = 0;
  for (int r = 0; r < nmeas; r++) {
    const orc_real br = (orc_real)b[r];
    const double* row = P->b2d + (size_t)r * n;
    for (int j = 0; j < n; j++) y_u[j] += (orc_real)row[j] * br;
  }
  icls_constraints(P, w, y_u, c_u);
  for (int r = num_ineq; r < ncon; r++) active[r] = 1;
  memcpy(c, c_u, sizeof(orc_real) * (size_t)ncon);
  memcpy(x, y_u, sizeof(orc_real) * (size_t)n);

  int niter = 0;
  for (;;) {
    if (num_ineq <= 0) break;
    int min_idx = 0;
    orc_real cmin = c[0];
    for (int r = 1; r < num_ineq; r++) if (c[r] < cmin) { cmin = c[r]; min_idx = r; }
    if (!(cmin < -(orc_real)P->tol)) break;
    int changed = !active[min_idx];
    active[min_idx] = 1;
    for (;;) {
      int na = 0;
      for (int r = 0; r < ncon; r++)
        if (active[r]) { act_idx[na] = r; l[na] = -c_u[r]; na++; }
      /* BtB = B_act B_act^T (lower): BtB_ij = sum_k B(a_i,k) B(a_j,k), k ascending */
      for (int j = 0; j < na; j++) {
        const double* bj = P->B + (size_t)act_idx[j] * n;
        for (int k = 0; k < n; k++) Gt[(size_t)k * na + j] = (orc_real)bj[k];
      }
      for (int i = 0; i < na; i++) {
        const double* bi = P->B + (size_t)act_idx[i] * n;
        orc_real* row = BtB + (size_t)i * na;
        for (int j = 0; j <= i; j++) row[j] = 0;
        for (int k = 0; k < n; k++) {
          const orc_real bik = (orc_real)bi[k];
          const orc_real* g = Gt + (size_t)k * na;
          for (int j = 0; j <= i; j++) row[j] += bik * g[j];
        }
      }
      for (int i = 0; i < na; i++) BtB[(size_t)i * na + i] += (orc_real)P->lambda_min_norm;
      llt_inplace(BtB, na);
      llt_solve_inplace(BtB, na, l);
      orc_real s_min = INFINITY;
      int s_min_index = 0;
      int a = 0;
      for (int r = 0; r < num_ineq; r++) {
        if (active[r]) {
          if (l[a] < 0.0) {
            orc_real sv = lambda_prev[r] / (lambda_prev[r] - l[a]);
            if (sv < s_min) { s_min = sv; s_min_index = r; }
          }
          lambda[r] = l[a];
          ++a;
        } else {
          lambda[r] = 0.0;
        }
      }
      if (!isfinite((double)s_min)) {
        /* y = y_u + B_act^T l: x_j = y_u_j + sum_i B(a_i,j) l_i (i ascending) */
        for (int j = 0; j < n; j++) x[j] = y_u[j];
        for (int i = 0; i < na; i++) {
          const double* bi = P->B + (size_t)act_idx[i] * n;
          const orc_real li = l[i];
          for (int j = 0; j < n; j++) x[j] += (orc_real)bi[j] * li;
        }
        break;
      }
      if (active[s_min_index]) changed = 1;
      active[s_min_index] = 0;
    }
    memcpy(lambda_prev, lambda, sizeof(orc_real) * (size_t)ncon);
    ++niter;
    if (!changed || niter > P->max_niter) break;
    icls_constraints(P, w, x, c);
  }
  /* x = L^-T y (constrained_least_squares.h:355) */
  for (int i = n - 1; i >= 0; i--) {
    orc_real s = x[i];
    for (int k = i + 1; k < n; k++) s -= (orc_real)P->chol[(size_t)k * n + i] * x[k];
    x[i] = s / (orc_real)P->chol[(size_t)i * n + i];
  }
  for (int j = 0; j < n; j++) xout[j] = (double)x[j];
  if (active_out) memcpy(active_out, active, (size_t)ncon);
  return niter;
}

int orc_icls_solve(const orc_icls_problem* P, const double* b, double* xout, uint8_t* active_out) {
  icls_ws w;
  icls_ws_alloc(&w, P);
  const int niter = icls_solve_ws(P, &w, b, xout, active_out);
  icls_ws_free(&w);
  return niter;
}

typedef struct icls_job {
  const orc_icls_problem* p; const float* b; size_t nvox; const uint8_t* valid;
  float* x; int16_t* iters; pthread_mutex_t mtx; size_t next;
} icls_job;

static void* icls_worker(void* arg) {
  icls_job* J = (icls_job*)arg;
  const int nm = J->p->nmeas, n = J->p->n;
  double* b = (double*)malloc(sizeof(double) * (size_t)nm);
  double* x = (double*)malloc(sizeof(double) * (size_t)n);
  icls_ws w;
  icls_ws_alloc(&w, J->p);
  const size_t chunk = 128;
  for (;;) {
    pthread_mutex_lock(&J->mtx);
    size_t v0 = J->next; J->next += chunk;
    pthread_mutex_unlock(&J->mtx);
    if (v0 >= J->nvox) break;
    size_t v1 = v0 + chunk < J->nvox ? v0 + chunk : J->nvox;
    for (size_t v = v0; v < v1; v++) {
      if (J->valid && !J->valid[v]) {
        for (int j = 0; j < n; j++) J->x[v * n + j] = 0.0f;
        if (J->iters) J->iters[v] = -1;
        continue;
      }
      for (int r = 0; r < nm; r++) b[r] = J->b[v * nm + r];
      int it = icls_solve_ws(J->p, &w, b, x, NULL);
      for (int j = 0; j < n; j++) J->x[v * n + j] = (float)x[j];
      if (J->iters) J->iters[v] = (int16_t)it;
    }
  }
  icls_ws_free(&w);
  free(b); free(x);
  return NULL;
}

int orc_icls_run_batch(const orc_icls_problem* p, const float* b, size_t nvox,
                       const uint8_t* valid, int nthreads, float* x, int16_t* iters) {
  icls_job J;
  J.p = p; J.b = b; J.nvox = nvox; J.valid = valid; J.x = x; J.iters = iters; J.next = 0;
  pthread_mutex_init(&J.mtx, NULL);
  if (nthreads <= 0) icls_worker(&J);
  else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, icls_worker, &J);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&J.mtx);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* MSMT_CSD::Shared::init matrices (msmt_csd.h:100-218)                                        */
/* ------------------------------------------------------------------------------------------ */
int orc_msmt_build(const double* grad, int nvol, const int* shell_of_vol, int nshell,
                   const double* responses, const int* resp_cols, int ntissue, const int* lmax,
                   const double* hr_dirs, int n_hr, double* C, double* A, int* ncols, int* nrows_A) {
  int nparams = 0, maxl = 0;
  for (int t = 0; t < ntissue; t++) { nparams += orc_nforl(lmax[t]); if (lmax[t] > maxl) maxl = lmax[t]; }
  const int nsh_max = orc_nforl(maxl);
  /* all-volume direction matrix; b=0 rows give NaN elevation -> SHT NaN -> 0 (:139-148) */
  double* dirs = (double*)malloc(sizeof(double) * 2 * (size_t)nvol);
  for (int v = 0; v < nvol; v++) {
    const double* g = grad + 4 * (size_t)v;
    dirs[2 * v] = atan2(g[1], g[0]);
    double z = g[2] / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (z >= 1.0) dirs[2 * v + 1] = 0.0;
    else if (z <= -1.0) dirs[2 * v + 1] = M_PI;
    else dirs[2 * v + 1] = acos(z);
  }
  orc_real* SHT = (orc_real*)malloc(sizeof(orc_real) * (size_t)nvol * nsh_max);
  sh_transform_r(dirs, nvol, maxl, SHT);
  for (size_t i = 0; i < (size_t)nvol * nsh_max; i++) if (isnan((double)SHT[i])) SHT[i] = 0.0;
  /* DSH: non-zero entries of the SH row at [0 0] = P_l^0(1) (:151-160) */
  orc_real dsh[64];
  {
    double d0[2] = {0.0, 0.0};
    orc_real* row = (orc_real*)malloc(sizeof(orc_real) * (size_t)nsh_max);
    sh_transform_r(d0, 1, maxl, row);
    int j = 0;
    for (int i = 0; i < nsh_max; i++) if (row[i] != 0.0) dsh[j++] = row[i];
    free(row);
  }
  for (size_t i = 0; i < (size_t)nvol * nparams; i++) C[i] = 0.0;
  int pbegin = 0, roff = 0;
  for (int t = 0; t < ntissue; t++) {
    const int tl = lmax[t], tn = orc_nforl(tl), tnz = tl / 2 + 1;
    const int rc = resp_cols[t];
    for (int s = 0; s < nshell; s++) {
      orc_real resp[64];
      for (int i = 0; i < tnz; i++) {
        orc_real r = i < rc ? (orc_real)responses[roff + (size_t)s * rc + i] : 0.0;
        resp[i] = r / dsh[i];
      }
      orc_real fconv[256];
      int li = 0, mi = 0;
      for (int l = 0; l <= tl; l += 2) {
        for (int m = -l; m <= l; m++) fconv[mi++] = resp[li];
        li++;
      }
      for (int v = 0; v < nvol; v++) {
        if (shell_of_vol[v] != s) continue;
        for (int k = 0; k < tn; k++)
          C[(size_t)v * nparams + pbegin + k] = (double)(SHT[(size_t)v * nsh_max + k] * fconv[k]);
      }
    }
    pbegin += tn;
    roff += nshell * rc;
  }
  /* A: block-diagonal, 300 rows for lmax>0 tissues, else 1x1 (:190-213) */
  orc_real* HR = (orc_real*)malloc(sizeof(orc_real) * (size_t)n_hr * nsh_max);
  sh_transform_r(hr_dirs, n_hr, maxl, HR);
  int Mr = 0;
  for (int t = 0; t < ntissue; t++) Mr += lmax[t] > 0 ? n_hr : 1;
  for (size_t i = 0; i < (size_t)Mr * nparams; i++) A[i] = 0.0;
  int bm = 0, bn = 0;
  for (int t = 0; t < ntissue; t++) {
    const int mt = lmax[t] > 0 ? n_hr : 1, nt = orc_nforl(lmax[t]);
    for (int i = 0; i < mt; i++)
      for (int j = 0; j < nt; j++) A[(size_t)(bm + i) * nparams + bn + j] = (double)HR[(size_t)i * nsh_max + j];
    bm += mt; bn += nt;
  }
  *ncols = nparams; *nrows_A = Mr;
  free(dirs); free(SHT); free(HR);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Single-thread timing of the two dense kernels of CSD::iterate (rank-k update and LLT)        */
/* ------------------------------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int orc_time_syrk_llt(const orc_csd_shared* s, int k, int reps, double* t_syrk, double* t_llt) {
  if (!s || k < 0 || k > s->n_hr || reps <= 0) return -1;
  csd_work w;
  work_alloc(&w, s);
  for (int q = 0; q < k; q++) w.neg[q] = (int)(((long)q * s->n_hr) / (k ? k : 1));
  w.nneg = k;
  double ts = 0, tl = 0, chk = 0;
  for (int r = 0; r < reps; r++) {
    const double a = now_s();
    csd_form_normal(s, &w);
    const double b = now_s();
    llt_inplace(w.W, s->n_sh);
    const double c = now_s();
    ts += b - a;
    tl += c - b;
    chk += (double)w.W[0];
  }
  work_free(&w);
  *t_syrk = ts;
  *t_llt = tl;
  return chk > 0 ? 0 : 1;
}
