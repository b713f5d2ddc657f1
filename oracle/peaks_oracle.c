/* CPU restatement of sh2peaks — TEST INFRASTRUCTURE ONLY (the checker of the GPU peak finder and
 * the CPU baseline of its benchmark; the product never links it).
 *
 * Follows the reference in its own arithmetic: single precision, with the double-precision steps
 * the reference takes where a double literal or constant enters an expression (marked (double)
 * below), and without fused multiply-adds (this file is compiled with -ffp-contract=off).
 *   Legendre::Plm_sph (array, lmax, m, x, sin_x)  core/math/legendre.h:129-154 (helper :66-69)
 *   Math::SH::derivatives                         core/math/SH.h:568-647
 *   Math::SH::get_peak                            core/math/SH.h:514-559 (constants :23-24)
 *   sh2peaks Processor / check_input / Direction  cmd/sh2peaks.cpp:93-105, 160-300
 *   -fast: Math::SH::PrecomputedAL (512 elevations) SH.h:402-470, Legendre::Plm_sph (x only)
 *          legendre.h:98-122, its use in derivatives() SH.h:589-594
 */
#define _GNU_SOURCE
#include "peaks_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MAX_DIR_CHANGE 0.2      /* SH.h:23 */
#define ANGLE_TOLERANCE 1e-4    /* SH.h:24 */
#define DOT_THRESHOLD 0.99      /* sh2peaks.cpp:25 */
#define SQRT2 1.41421356237309504880   /* Math::sqrt2 (core/math/math.h:43), a double */

static int sh_index(int l, int m) { return l * (l + 1) / 2 + m; }     /* SH.h:52-55 */
static int sh_index_mpos(int l, int m) { return l * l / 4 + m; }     /* SH.h:64-67 */

/* legendre.h:66-69, called there with double arguments */
static double plm_helper(double x, double m) { return m < 1.0 ? 1.0 : x * (m - 1.0) / m * plm_helper(x, m - 2.0); }

/* legendre.h:129-154: P_l^m for l = m..lmax into array[l], from cos and sin of the elevation */
void orc_plm_sph_f(float* array, int lmax, int m, float x, float sin_x) {
  const float sin2 = sin_x * sin_x;
  if (m && sin2 == 0.0f) {
    for (int n = m; n <= lmax; ++n) array[n] = 0.0f;
    return;
  }
  array[m] = (float)0.282094791773878;
  if (m) array[m] = (float)((double)array[m] * sqrt((double)(float)(2 * m + 1) * plm_helper(1.0 * sin2, 2.0 * m)));
  if (m & 1) array[m] = -array[m];
  if (lmax == m) return;
  float f = sqrtf((float)(2 * m + 3));
  array[m + 1] = x * f * array[m];
  for (int n = m + 2; n <= lmax; n++) {
    array[n] = x * array[n - 1] - array[n - 2] / f;
    f = sqrtf((float)(4 * n * n - 1) / (float)(n * n - m * m));
    array[n] *= f;
  }
}

/* legendre.h:98-122: the same recursion from cos(elevation) alone (sin^2 as 1 - x^2) */
static void plm_sph_x_f(float* array, int lmax, int m, float x) {
  const float x2 = x * x;
  if (m && (double)x2 >= 1.0) {
    for (int n = m; n <= lmax; ++n) array[n] = 0.0f;
    return;
  }
  array[m] = (float)0.282094791773878;
  if (m) array[m] = (float)((double)array[m] * sqrt((double)(float)(2 * m + 1) * plm_helper(1.0 - (double)x2, 2.0 * m)));
  if (m & 1) array[m] = -array[m];
  if (lmax == m) return;
  float f = sqrtf((float)(2 * m + 3));
  array[m + 1] = x * f * array[m];
  for (int n = m + 2; n <= lmax; n++) {
    array[n] = x * array[n - 1] - array[n - 2] / f;
    f = sqrtf((float)(4 * n * n - 1) / (float)(n * n - m * m));
    array[n] *= f;
  }
}

/* SH.h:418-434 (PrecomputedAL::init, 512 directions as sh2peaks constructs it): table
 * [ORC_PEAKS_FAST_NDIR][(lmax/2+1)^2], row n the values at elevation n * pi / 511.  Returns the
 * row length. */
int orc_peaks_fast_table(int lmax, float* table) {
  const int nAL = (lmax / 2 + 1) * (lmax / 2 + 1);
  const float inc = (float)(M_PI / (ORC_PEAKS_FAST_NDIR - 1));
  float buf[ORC_PEAKS_MAX_LMAX + 1];
  for (int n = 0; n < ORC_PEAKS_FAST_NDIR; n++) {
    float* p = table + (size_t)n * nAL;
    const float cos_el = cosf((float)n * inc);
    for (int m = 0; m <= lmax; m++) {
      plm_sph_x_f(buf, lmax, m, cos_el);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) p[sh_index_mpos(l, m)] = buf[l];
    }
  }
  return nAL;
}

/* SH.h:436-470 (set / get): linear interpolation between the two rows around the elevation,
 * the first or last row alone outside [0, pi) */
static void fast_al(const float* table, int lmax, float elevation, float* AL) {
  const int nAL = (lmax / 2 + 1) * (lmax / 2 + 1);
  const float inc = (float)(M_PI / (ORC_PEAKS_FAST_NDIR - 1));
  float f2 = elevation / inc, f1;
  int i = (int)f2;
  if (i < 0) {
    i = 0; f1 = 1.0f; f2 = 0.0f;
  } else if (i >= ORC_PEAKS_FAST_NDIR - 1) {
    i = ORC_PEAKS_FAST_NDIR - 1; f1 = 1.0f; f2 = 0.0f;
  } else {
    f2 -= (float)i;
    f1 = (float)(1.0 - (double)f2);
  }
  const float* p1 = table + (size_t)i * nAL;
  const float* p2 = p1 + nAL;
  for (int l = 0; l <= lmax; l += 2)
    for (int m = 0; m <= l; m++) {
      const int k = sh_index_mpos(l, m);
      float v = f1 * p1[k];
      if (f2 != 0.0f) v += f2 * p2[k];
      AL[k] = v;
    }
}

/* SH.h:568-647: amplitude and first/second derivatives in elevation/azimuth.
 * out = {amplitude, dSH_del, dSH_daz, d2SH_del2, d2SH_deldaz, d2SH_daz2} */
static void derivatives_impl(const float* sh, int lmax, float elevation, float azimuth, const float* table,
                             float* out) {
  const float sel = sinf(elevation);
  const float cel = cosf(elevation);
  const int atpole = (double)sel < 1e-4;
  float dSH_del = 0.0f, dSH_daz = 0.0f, d2SH_del2 = 0.0f, d2SH_deldaz = 0.0f, d2SH_daz2 = 0.0f;
  float AL[64] = {0.0f}, buf[ORC_PEAKS_MAX_LMAX + 1];   /* NforL_mpos(14) = 64 */
  if (lmax < 0 || lmax > ORC_PEAKS_MAX_LMAX || (lmax & 1)) {
    for (int i = 0; i < 6; i++) out[i] = NAN;
    return;
  }
  if (table) {
    fast_al(table, lmax, elevation, AL);
  } else {
    for (int m = 0; m <= lmax; m++) {
      orc_plm_sph_f(buf, lmax, m, cel, sel);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) AL[sh_index_mpos(l, m)] = buf[l];
    }
  }
  float amplitude = sh[0] * AL[0];
  for (int l = 2; l <= lmax; l += 2) {
    const float v = sh[sh_index(l, 0)];
    amplitude += v * AL[sh_index_mpos(l, 0)];
    dSH_del += v * sqrtf((float)(l * (l + 1))) * AL[sh_index_mpos(l, 1)];
    const float t = v * (sqrtf((float)(l * (l + 1) * (l - 1) * (l + 2))) * AL[sh_index_mpos(l, 2)] -
                         (float)(l * (l + 1)) * AL[sh_index_mpos(l, 0)]);
    d2SH_del2 = (float)((double)d2SH_del2 + (double)t / 2.0);
  }
  for (int m = 1; m <= lmax; m++) {
    const float caz = (float)(SQRT2 * (double)cosf((float)m * azimuth));
    const float saz = (float)(SQRT2 * (double)sinf((float)m * azimuth));
    for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) {
      const float vp = sh[sh_index(l, m)];
      const float vm = sh[sh_index(l, -m)];
      const float al = AL[sh_index_mpos(l, m)];
      amplitude += (vp * caz + vm * saz) * al;

      float tmp = sqrtf((float)((l + m) * (l - m + 1))) * AL[sh_index_mpos(l, m - 1)];
      if (l > m) tmp -= sqrtf((float)((l - m) * (l + m + 1))) * AL[sh_index_mpos(l, m + 1)];
      tmp = (float)((double)tmp / -2.0);
      dSH_del += (vp * caz + vm * saz) * tmp;

      float tmp2 = (float)(-((l + m) * (l - m + 1) + (l - m) * (l + m + 1))) * al;
      if (m == 1)
        tmp2 -= sqrtf((float)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[sh_index_mpos(l, 1)];
      else
        tmp2 += sqrtf((float)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[sh_index_mpos(l, m - 2)];
      if (l > m + 1)
        tmp2 += sqrtf((float)((l - m) * (l + m + 1) * (l - m - 1) * (l + m + 2))) * AL[sh_index_mpos(l, m + 2)];
      tmp2 = (float)((double)tmp2 / 4.0);
      d2SH_del2 += (vp * caz + vm * saz) * tmp2;

      if (atpole) {
        dSH_daz += (vm * caz - vp * saz) * tmp;
      } else {
        d2SH_deldaz += (float)m * (vm * caz - vp * saz) * tmp;
        dSH_daz += (float)m * (vm * caz - vp * saz) * al;
        d2SH_daz2 -= (vp * caz + vm * saz) * (float)m * (float)m * al;
      }
    }
  }
  if (!atpole) {
    dSH_daz /= sel;
    d2SH_deldaz /= sel;
    d2SH_daz2 /= sel * sel;
  }
  out[0] = amplitude;
  out[1] = dSH_del;
  out[2] = dSH_daz;
  out[3] = d2SH_del2;
  out[4] = d2SH_deldaz;
  out[5] = d2SH_daz2;
}

void orc_sh_derivatives_f(const float* sh, int lmax, float elevation, float azimuth, float* out) {
  derivatives_impl(sh, lmax, elevation, azimuth, NULL, out);
}

/* Eigen's normalize(): divide by the norm when the squared norm is positive */
static void normalize3(float* v) {
  const float z = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  if (z > 0.0f) {
    const float s = sqrtf(z);
    v[0] /= s;
    v[1] /= s;
    v[2] /= s;
  }
}

/* SH.h:514-559: Gauss-Newton ascent from dir (unit, updated in place); returns the amplitude, or
 * NaN (and a NaN direction) after 50 iterations without convergence.  *iters gets the count. */
static float get_peak_impl(const float* sh, int lmax, float* dir, int* iters, const float* table) {
  for (int i = 0; i < 50; i++) {
    const float az = atan2f(dir[1], dir[0]);
    const float el = atan2f(hypotf(dir[0], dir[1]), dir[2]);
    float d[6];
    derivatives_impl(sh, lmax, el, az, table, d);
    const float amplitude = d[0], dSH_del = d[1], dSH_daz = d[2], d2SH_del2 = d[3], d2SH_deldaz = d[4],
                d2SH_daz2 = d[5];
    float del = sqrtf(dSH_del * dSH_del + dSH_daz * dSH_daz);
    float daz = 0.0f;
    if (del != 0.0f) {
      daz = dSH_daz / del;
      del = dSH_del / del;
    }
    const float dSH_dt = daz * dSH_daz + del * dSH_del;
    const float d2SH_dt2 = (float)((double)(daz * daz * d2SH_daz2) + 2.0 * (double)daz * (double)del * (double)d2SH_deldaz +
                                   (double)(del * del * d2SH_del2));
    float dt = d2SH_dt2 != 0.0f ? (-dSH_dt / d2SH_dt2) : (float)MAX_DIR_CHANGE;
    if (dt < 0.0f) dt = -dt;
    if ((double)dt > MAX_DIR_CHANGE) dt = (float)MAX_DIR_CHANGE;
    del *= dt;
    daz *= dt;
    dir[0] += del * cosf(az) * cosf(el) - daz * sinf(az);
    dir[1] += del * sinf(az) * cosf(el) + daz * cosf(az);
    dir[2] -= del * sinf(el);
    normalize3(dir);
    if ((double)dt < ANGLE_TOLERANCE) {
      if (iters) *iters = i + 1;
      return amplitude;
    }
  }
  dir[0] = dir[1] = dir[2] = NAN;
  if (iters) *iters = 50;
  return NAN;
}

float orc_sh_get_peak_f(const float* sh, int lmax, float* dir, int* iters) {
  return get_peak_impl(sh, lmax, dir, iters, NULL);
}

static float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* sh2peaks.cpp:266-286 (check_input): a NaN coefficient, or no non-zero coefficient past the
 * first, gives an all-NaN voxel */
static int no_peaks_input(const float* sh, int nsh) {
  int none = 1;
  for (int i = 0; i < nsh; i++) {
    if (isnan(sh[i])) return 1;
    if (none && i && sh[i] != 0.0f) none = 0;
  }
  return none;
}

/* sh2peaks.cpp:160-247: one voxel.  seeds [nseeds][2] (azimuth, elevation); match [npeaks][3]
 * (nullable): directions to report the closest peak to (-direction, or a voxel of -peaks), each
 * normalised here; out [3 npeaks] = amplitude * direction of each reported peak, NaN past the
 * last.  Returns the number of accepted peaks. */
static int voxel_impl(const float* sh, int nsh, int lmax, const float* seeds, int nseeds, int npeaks,
                      float threshold, const float* match, float* out, int* total_iters, const float* table) {
  for (int i = 0; i < 3 * npeaks; i++) out[i] = NAN;
  if (no_peaks_input(sh, nsh)) return 0;
  if (nseeds > ORC_PEAKS_MAX_SEEDS) nseeds = ORC_PEAKS_MAX_SEEDS;
  float pa[ORC_PEAKS_MAX_SEEDS], pv[ORC_PEAKS_MAX_SEEDS][3];
  int np = 0;
  for (int s = 0; s < nseeds; s++) {
    const float phi = seeds[2 * s], theta = seeds[2 * s + 1];
    float v[3] = {cosf(phi) * sinf(theta), sinf(phi) * sinf(theta), cosf(theta)};   /* sh2peaks.cpp:96 */
    int it = 0;
    float a = get_peak_impl(sh, lmax, v, &it, table);
    if (total_iters) *total_iters += it;
    if (isfinite(a)) {
      for (int j = 0; j < np; j++)
        if ((double)fabsf(dot3(v, pv[j])) > DOT_THRESHOLD) {
          a = NAN;
          break;
        }
    }
    if (isfinite(a) && a >= threshold) {
      pa[np] = a;
      memcpy(pv[np], v, sizeof v);
      np++;
    }
  }
  int sel[ORC_PEAKS_MAX_SEEDS];
  int nout = npeaks < np ? npeaks : np;
  if (match) {
    /* the accepted peak most nearly parallel to each given direction (sh2peaks.cpp:205-235);
     * none with |dot| > 0 leaves that slot NaN */
    for (int i = 0; i < npeaks; i++) {
      float p[3] = {match[3 * i], match[3 * i + 1], match[3 * i + 2]};
      normalize3(p);
      float mdot = 0.0f;
      sel[i] = -1;
      for (int n = 0; n < np; n++) {
        const float f = fabsf(dot3(p, pv[n]));
        if (f > mdot) {
          mdot = f;
          sel[i] = n;
        }
      }
    }
  } else {
    /* partial_sort_copy by decreasing amplitude (sh2peaks.cpp:237); equal amplitudes keep seed
     * order */
    int used[ORC_PEAKS_MAX_SEEDS] = {0};
    for (int i = 0; i < nout; i++) {
      int best = -1;
      for (int n = 0; n < np; n++)
        if (!used[n] && (best < 0 || pa[n] > pa[best])) best = n;
      used[best] = 1;
      sel[i] = best;
    }
  }
  for (int i = 0; i < nout; i++) {
    const int n = sel[i];
    if (n < 0) continue;
    out[3 * i] = pa[n] * pv[n][0];
    out[3 * i + 1] = pa[n] * pv[n][1];
    out[3 * i + 2] = pa[n] * pv[n][2];
  }
  return np;
}

int orc_sh2peaks_voxel(const float* sh, int nsh, int lmax, const float* seeds, int nseeds, int npeaks,
                       float threshold, const float* match, float* out, int* total_iters) {
  return voxel_impl(sh, nsh, lmax, seeds, nseeds, npeaks, threshold, match, out, total_iters, NULL);
}

typedef struct peaks_job {
  const float* sh;
  int ld, nsh, lmax, nseeds, npeaks, ld_match;
  const float* table;   /* -fast lookup table or NULL */
  const float* seeds;
  float threshold;
  const float* match;
  const uint8_t* mask;
  float* out;
  size_t nvox, next;
  long long iters;
  pthread_mutex_t mtx;
} peaks_job;

static void* peaks_worker(void* arg) {
  peaks_job* J = (peaks_job*)arg;
  const size_t chunk = 64;
  long long iters = 0;
  for (;;) {
    pthread_mutex_lock(&J->mtx);
    const size_t v0 = J->next;
    J->next += chunk;
    pthread_mutex_unlock(&J->mtx);
    if (v0 >= J->nvox) break;
    const size_t v1 = v0 + chunk < J->nvox ? v0 + chunk : J->nvox;
    for (size_t v = v0; v < v1; v++) {
      float* o = J->out + v * (size_t)(3 * J->npeaks);
      if (J->mask && !J->mask[v]) {
        for (int i = 0; i < 3 * J->npeaks; i++) o[i] = NAN;
        continue;
      }
      const float* m = J->match ? J->match + v * (size_t)J->ld_match : NULL;
      if (m && isnan(m[0])) {   /* sh2peaks.cpp:268-274 */
        for (int i = 0; i < 3 * J->npeaks; i++) o[i] = NAN;
        continue;
      }
      int it = 0;
      voxel_impl(J->sh + v * (size_t)J->ld, J->nsh, J->lmax, J->seeds, J->nseeds, J->npeaks, J->threshold, m, o, &it,
                 J->table);
      iters += it;
    }
  }
  pthread_mutex_lock(&J->mtx);
  J->iters += iters;
  pthread_mutex_unlock(&J->mtx);
  return NULL;
}

/* The whole volume on nthreads workers (0: the calling thread).  sh [nvox][ld] (the first nsh
 * values of a row are the coefficients); match [nvox][ld_match] per voxel (ld_match 0: the same
 * directions for every voxel) or NULL; mask [nvox] or NULL.  Returns the Newton iterations run. */
long long orc_sh2peaks_run_fast(const float* sh, int ld, int nsh, size_t nvox, int lmax, const float* seeds, int nseeds,
                                int npeaks, float threshold, const float* match, int ld_match, const uint8_t* mask,
                                int nthreads, const float* table, float* out) {
  peaks_job J;
  J.table = table;
  J.sh = sh; J.ld = ld; J.nsh = nsh; J.lmax = lmax; J.seeds = seeds; J.nseeds = nseeds; J.npeaks = npeaks;
  J.threshold = threshold; J.match = match; J.ld_match = ld_match; J.mask = mask; J.out = out;
  J.nvox = nvox; J.next = 0; J.iters = 0;
  pthread_mutex_init(&J.mtx, NULL);
  if (nthreads <= 0) {
    peaks_worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, peaks_worker, &J);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&J.mtx);
  return J.iters;
}

long long orc_sh2peaks_run(const float* sh, int ld, int nsh, size_t nvox, int lmax, const float* seeds, int nseeds,
                           int npeaks, float threshold, const float* match, int ld_match, const uint8_t* mask,
                           int nthreads, float* out) {
  return orc_sh2peaks_run_fast(sh, ld, nsh, nvox, lmax, seeds, nseeds, npeaks, threshold, match, ld_match, mask,
                               nthreads, NULL, out);
}
