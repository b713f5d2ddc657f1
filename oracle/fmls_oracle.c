/* CPU oracle of FOD segmentation (fod2fixel's FMLS) -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of MRtrix3's
 *   - DWI::FMLS::Segmenter::operator()           (src/dwi/fmls.cpp:166-377)
 *   - FOD_lobe add / merge / finalise             (src/dwi/fmls.h:61-137)
 *   - DWI::Directions::FastLookupSet::select_direction and its grid
 *                                                 (src/dwi/directions/set.cpp:336-467)
 *   - Math::SH::PrecomputedAL<double>, derivatives and get_peak with it
 *                                                 (core/math/SH.h:398-647)
 * used by tests/ as the checker of the GPU segmentation.  The direction adjacency and the
 * integration weights are inputs (the tests compute them independently: a convex hull and a
 * least-squares fit), so this file follows the segmentation itself step by step, including the
 * order in which directions are visited, lobe merges and retrospective assignments.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FMLS_MAX_ADJ 8
#define FMLS_MAX_LOBES 1281
#define FMLS_MAX_PEAKS 16
#define FMLS_MAX_LMAX 16

static int nforl(int l) { return (l + 1) * (l + 2) / 2; }
static int sh_index(int l, int m) { return l * (l + 1) / 2 + m; }
static int sh_index_mpos(int l, int m) { return l * l / 4 + m; }

/* ---- Legendre::Plm_sph, one-argument form (core/math/legendre.h:98-122), double ---- */
static double plm_helper(double x, double m) { return m < 1.0 ? 1.0 : x * (m - 1.0) / m * plm_helper(x, m - 2.0); }

static void plm_sph(double* array, int lmax, int m, double x) {
  const double x2 = x * x;
  if (m && x2 >= 1.0) {
    for (int n = m; n <= lmax; ++n) array[n] = 0.0;
    return;
  }
  array[m] = 0.282094791773878;
  if (m) array[m] *= sqrt((double)(2 * m + 1) * plm_helper(1.0 - x2, 2.0 * m));
  if (m & 1) array[m] = -array[m];
  if (lmax == m) return;
  double f = sqrt((double)(2 * m + 3));
  array[m + 1] = x * f * array[m];
  for (int n = m + 2; n <= lmax; n++) {
    array[n] = x * array[n - 1] - array[n - 2] / f;
    f = sqrt((double)(4 * n * n - 1) / (double)(n * n - m * m));
    array[n] *= f;
  }
}

/* PrecomputedAL<double>::init (SH.h:418-434): table [ndir][(lmax/2+1)^2], row n at elevation
 * n pi / (ndir - 1).  Returns the row length. */
int orc_fmls_al_table(int lmax, int ndir, double* table) {
  const int nAL = (lmax / 2 + 1) * (lmax / 2 + 1);
  const double inc = M_PI / (ndir - 1);
  double buf[FMLS_MAX_LMAX + 1];
  for (int n = 0; n < ndir; n++) {
    double* p = table + (size_t)n * nAL;
    const double cos_el = cos(n * inc);
    for (int m = 0; m <= lmax; m++) {
      plm_sph(buf, lmax, m, cos_el);
      for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) p[sh_index_mpos(l, m)] = buf[l];
    }
  }
  return nAL;
}

/* set / get (SH.h:436-470) */
static void al_interp(const double* table, int ndir, int lmax, double elevation, double* AL) {
  const int nAL = (lmax / 2 + 1) * (lmax / 2 + 1);
  const double inc = M_PI / (ndir - 1);
  double f2 = elevation / inc, f1;
  int i = (int)f2;
  if (i < 0) {
    i = 0; f1 = 1.0; f2 = 0.0;
  } else if (i >= ndir - 1) {
    i = ndir - 1; f1 = 1.0; f2 = 0.0;
  } else {
    f2 -= i;
    f1 = 1.0 - f2;
  }
  const double* p1 = table + (size_t)i * nAL;
  const double* p2 = p1 + nAL;
  for (int l = 0; l <= lmax; l += 2)
    for (int m = 0; m <= l; m++) {
      const int k = sh_index_mpos(l, m);
      double v = f1 * p1[k];
      if (f2) v += f2 * p2[k];
      AL[k] = v;
    }
}

/* derivatives (SH.h:568-647) with the precomputed table; out = {amplitude, dSH_del, dSH_daz,
 * d2SH_del2, d2SH_deldaz, d2SH_daz2} */
static void derivatives(const double* sh, int lmax, double elevation, double azimuth, const double* table, int ndir,
                        double* out) {
  const double sel = sin(elevation);
  const int atpole = sel < 1e-4;
  double dSH_del = 0.0, dSH_daz = 0.0, d2SH_del2 = 0.0, d2SH_deldaz = 0.0, d2SH_daz2 = 0.0;
  double AL[81] = {0.0};
  al_interp(table, ndir, lmax, elevation, AL);
  double amplitude = sh[0] * AL[0];
  for (int l = 2; l <= lmax; l += 2) {
    const double v = sh[sh_index(l, 0)];
    amplitude += v * AL[sh_index_mpos(l, 0)];
    dSH_del += v * sqrt((double)(l * (l + 1))) * AL[sh_index_mpos(l, 1)];
    d2SH_del2 += v * (sqrt((double)(l * (l + 1) * (l - 1) * (l + 2))) * AL[sh_index_mpos(l, 2)] -
                      l * (l + 1) * AL[sh_index_mpos(l, 0)]) / 2.0;
  }
  for (int m = 1; m <= lmax; m++) {
    const double caz = M_SQRT2 * cos(m * azimuth);
    const double saz = M_SQRT2 * sin(m * azimuth);
    for (int l = (m & 1) ? m + 1 : m; l <= lmax; l += 2) {
      const double vp = sh[sh_index(l, m)];
      const double vm = sh[sh_index(l, -m)];
      amplitude += (vp * caz + vm * saz) * AL[sh_index_mpos(l, m)];
      double tmp = sqrt((double)((l + m) * (l - m + 1))) * AL[sh_index_mpos(l, m - 1)];
      if (l > m) tmp -= sqrt((double)((l - m) * (l + m + 1))) * AL[sh_index_mpos(l, m + 1)];
      tmp /= -2.0;
      dSH_del += (vp * caz + vm * saz) * tmp;
      double tmp2 = -((l + m) * (l - m + 1) + (l - m) * (l + m + 1)) * AL[sh_index_mpos(l, m)];
      if (m == 1) tmp2 -= sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[sh_index_mpos(l, 1)];
      else tmp2 += sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[sh_index_mpos(l, m - 2)];
      if (l > m + 1) tmp2 += sqrt((double)((l - m) * (l + m + 1) * (l - m - 1) * (l + m + 2))) * AL[sh_index_mpos(l, m + 2)];
      tmp2 /= 4.0;
      d2SH_del2 += (vp * caz + vm * saz) * tmp2;
      if (atpole) {
        dSH_daz += (vm * caz - vp * saz) * tmp;
      } else {
        d2SH_deldaz += m * (vm * caz - vp * saz) * tmp;
        dSH_daz += m * (vm * caz - vp * saz) * AL[sh_index_mpos(l, m)];
        d2SH_daz2 -= (vp * caz + vm * saz) * m * m * AL[sh_index_mpos(l, m)];
      }
    }
  }
  if (!atpole) {
    dSH_daz /= sel;
    d2SH_deldaz /= sel;
    d2SH_daz2 /= sel * sel;
  }
  out[0] = amplitude; out[1] = dSH_del; out[2] = dSH_daz;
  out[3] = d2SH_del2; out[4] = d2SH_deldaz; out[5] = d2SH_daz2;
}

/* get_peak (SH.h:514-559) in double with the precomputed table; dir updated in place, NaN when
 * it does not converge in 50 steps */
double orc_fmls_get_peak(const double* sh, int lmax, double* dir, const double* table, int ndir) {
  const double MAX_DIR_CHANGE = 0.2, ANGLE_TOLERANCE = 1e-4;
  for (int i = 0; i < 50; i++) {
    const double az = atan2(dir[1], dir[0]);
    const double el = atan2(hypot(dir[0], dir[1]), dir[2]);
    double d[6];
    derivatives(sh, lmax, el, az, table, ndir, d);
    double del = sqrt(d[1] * d[1] + d[2] * d[2]);
    double daz = 0.0;
    if (del != 0.0) {
      daz = d[2] / del;
      del = d[1] / del;
    }
    const double dSH_dt = daz * d[2] + del * d[1];
    const double d2SH_dt2 = daz * daz * d[5] + 2.0 * daz * del * d[4] + del * del * d[3];
    double dt = d2SH_dt2 ? (-dSH_dt / d2SH_dt2) : MAX_DIR_CHANGE;
    if (dt < 0.0) dt = -dt;
    if (dt > MAX_DIR_CHANGE) dt = MAX_DIR_CHANGE;
    del *= dt;
    daz *= dt;
    dir[0] += del * cos(az) * cos(el) - daz * sin(az);
    dir[1] += del * sin(az) * cos(el) + daz * cos(az);
    dir[2] -= del * sin(el);
    const double nrm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] /= nrm; dir[1] /= nrm; dir[2] /= nrm;
    if (dt < ANGLE_TOLERANCE) return d[0];
  }
  dir[0] = dir[1] = dir[2] = NAN;
  return NAN;
}

/* ---- FastLookupSet (set.cpp:336-467): a grid over (azimuth, elevation) whose cells list the
 * directions that can be nearest to a point of the cell ---- */
typedef struct orc_lookup {
  int n_az, n_el;
  double az_step, el_step, az_begin, el_begin;
  int* start;   /* [n_az * n_el + 1] offsets into idx */
  int* idx;
} orc_lookup;

static int select_slow(const double* unit, int n, const double* p) {
  int best = 0;
  double mx = fabs(p[0] * unit[0] + p[1] * unit[1] + p[2] * unit[2]);
  for (int i = 1; i < n; i++) {
    const double dp = fabs(p[0] * unit[3 * i] + p[1] * unit[3 * i + 1] + p[2] * unit[3 * i + 2]);
    if (dp > mx) { mx = dp; best = i; }
  }
  return best;
}

static size_t dir2grid(const orc_lookup* L, const double* p) {
  const double az = atan2(p[1], p[0]);
  const double el = acos(p[2]);
  const size_t ia = (size_t)floor((az - L->az_begin) / L->az_step);
  const size_t ie = (size_t)floor((el - L->el_begin) / L->el_step);
  return ia * (size_t)L->n_el + ie;
}

static int cmp_int(const void* a, const void* b) { return *(const int*)a - *(const int*)b; }

/* builds the grid; adj [n][FMLS_MAX_ADJ] (-1 padded).  Returns the lookup (free with
 * orc_fmls_lookup_free). */
orc_lookup* orc_fmls_lookup_create(const double* unit, int n, const int* adj) {
  double sum = 0.0;
  size_t cnt = 0;
  for (int i = 0; i < n; i++)
    for (int a = 0; a < FMLS_MAX_ADJ; a++) {
      const int j = adj[i * FMLS_MAX_ADJ + a];
      if (j > i) {
        sum += fabs(unit[3 * i] * unit[3 * j] + unit[3 * i + 1] * unit[3 * j + 1] + unit[3 * i + 2] * unit[3 * j + 2]);
        cnt++;
      }
    }
  const double min_dp = sum / (double)cnt;
  const double max_angle_step = acos(min_dp);
  orc_lookup* L = (orc_lookup*)calloc(1, sizeof(orc_lookup));
  L->n_az = (int)ceil(2.0 * M_PI / max_angle_step);
  L->n_el = (int)ceil(M_PI / max_angle_step);
  const int total = L->n_az * L->n_el;
  L->az_step = 2.0 * M_PI / (double)(L->n_az - 1);
  L->el_step = M_PI / (double)(L->n_el - 1);
  L->az_begin = -M_PI;
  L->el_begin = 0.0;
  /* cell lists with generous capacity */
  const int cap = 64;
  int* lists = (int*)malloc(sizeof(int) * (size_t)total * cap);
  int* cnts = (int*)calloc((size_t)total, sizeof(int));
  for (int i = 0; i < n; i++) {
    const size_t g = dir2grid(L, unit + 3 * i);
    lists[g * cap + cnts[g]++] = i;
  }
  for (int g = 0; g < total; g++) {
    const int ia = g / L->n_el, ie = g - ia * L->n_el;
    for (int pidx = 0; pidx < 4; pidx++) {
      double az = L->az_begin + ia * L->az_step, el = L->el_begin + ie * L->el_step;
      if (pidx == 1) az += L->az_step;
      else if (pidx == 2) { az += L->az_step; el += L->el_step; }
      else if (pidx == 3) el += L->el_step;
      const double p[3] = {cos(az) * sin(el), sin(az) * sin(el), cos(el)};
      const int nd = select_slow(unit, n, p);
      int present = 0;
      for (int k = 0; k < cnts[g] && !present; k++) present = lists[g * cap + k] == nd;
      if (!present) lists[g * cap + cnts[g]++] = nd;
    }
  }
  for (int g = 0; g < total; g++) {
    const int nexp = cnts[g];
    for (int e = 0; e < nexp; e++) {
      const int d = lists[g * cap + e];
      for (int a = 0; a < FMLS_MAX_ADJ; a++) {
        const int j = adj[d * FMLS_MAX_ADJ + a];
        if (j < 0) continue;
        int present = 0;
        for (int k = 0; k < cnts[g] && !present; k++) present = lists[g * cap + k] == j;
        if (!present) lists[g * cap + cnts[g]++] = j;
      }
    }
    qsort(lists + (size_t)g * cap, (size_t)cnts[g], sizeof(int), cmp_int);
  }
  L->start = (int*)malloc(sizeof(int) * (size_t)(total + 1));
  L->start[0] = 0;
  for (int g = 0; g < total; g++) L->start[g + 1] = L->start[g] + cnts[g];
  L->idx = (int*)malloc(sizeof(int) * (size_t)L->start[total]);
  for (int g = 0; g < total; g++) memcpy(L->idx + L->start[g], lists + (size_t)g * cap, sizeof(int) * (size_t)cnts[g]);
  free(lists);
  free(cnts);
  return L;
}

void orc_fmls_lookup_free(orc_lookup* L) {
  if (!L) return;
  free(L->start);
  free(L->idx);
  free(L);
}

/* grid shape and the cell lists, for comparison with the product's tables */
int orc_fmls_lookup_info(const orc_lookup* L, int* n_az, int* n_el, int* start, int* idx, int cap) {
  const int total = L->n_az * L->n_el;
  *n_az = L->n_az;
  *n_el = L->n_el;
  if (start) memcpy(start, L->start, sizeof(int) * (size_t)(total + 1));
  if (idx && cap >= L->start[total]) memcpy(idx, L->idx, sizeof(int) * (size_t)L->start[total]);
  return L->start[total];
}

int orc_fmls_select(const orc_lookup* L, const double* unit, const double* p) {
  const size_t g = dir2grid(L, p);
  int best = L->idx[L->start[g]];
  double mx = fabs(p[0] * unit[3 * best] + p[1] * unit[3 * best + 1] + p[2] * unit[3 * best + 2]);
  for (int k = L->start[g] + 1; k < L->start[g + 1]; k++) {
    const int d = L->idx[k];
    const double dp = fabs(p[0] * unit[3 * d] + p[1] * unit[3 * d + 1] + p[2] * unit[3 * d + 2]);
    if (dp > mx) { mx = dp; best = d; }
  }
  return best;
}

/* ---- the segmentation (fmls.cpp:166-377) ---- */
typedef struct lobe {
  int neg;
  double max_peak;
  int npeaks;
  double peaks[FMLS_MAX_PEAKS][3];
  double mean[3];
  double integral;
} lobe;

typedef struct vo_pair { double v; int i; } vo_pair;

/* multimap<value, index, Max_abs>: |value| descending, insertion (index) order among equal keys */
static int cmp_absdesc(const void* a, const void* b) {
  const vo_pair* x = (const vo_pair*)a;
  const vo_pair* y = (const vo_pair*)b;
  const double ax = fabs(x->v), ay = fabs(y->v);
  if (ax > ay) return -1;
  if (ax < ay) return 1;
  return x->i - y->i;
}

static void lobe_add(lobe* L, const double* unit, int d, double value, double weight) {
  const double* dir = unit + 3 * d;
  const double mult = (L->mean[0] * dir[0] + L->mean[1] * dir[1] + L->mean[2] * dir[2]) > 0.0 ? 1.0 : -1.0;
  for (int c = 0; c < 3; c++) L->mean[c] += dir[c] * mult * fabs(value) * weight;
  L->integral += fabs(value * weight);
}

/* Segments one FOD.  SHT [ndir][nforl(lmax)] is the amplitude transform of the directions,
 * unit [ndir][3] their unit vectors, adj [ndir][FMLS_MAX_ADJ] (-1 padded) their neighbours,
 * weights [ndir] the integration weights; al_table [al_ndir][..] from orc_fmls_al_table (NULL:
 * the Newton revision of the peaks is skipped, for structural tests on synthetic direction graphs).
 * Output per lobe (in the reference's order): mean_dir[3], peak_dir[3] (of the maximal peak),
 * integral, max_peak -> out [max_out][8]; lut [ndir] (lobe of each direction, nlobes if none).
 * Returns the number of lobes (capped at max_out in the output, the count returned in full), or
 * -1 when more than FMLS_MAX_LOBES lobes arise before thresholding. */
int orc_fmls_segment(const double* sh_in, int lmax, const double* SHT, int ndir, const double* unit, const int* adj,
                     const double* weights, const orc_lookup* look, const double* al_table, int al_ndir,
                     double integral_threshold, double peak_value_threshold, double merge_ratio,
                     double* out, int max_out, int* lut) {
  const int nsh = nforl(lmax);
  if (!(sh_in[0] > 0.0) || !isfinite(sh_in[0])) {
    if (lut) for (int d = 0; d < ndir; d++) lut[d] = 0;
    return 0;
  }
  double* values = (double*)malloc(sizeof(double) * (size_t)ndir);
  vo_pair* order = (vo_pair*)malloc(sizeof(vo_pair) * (size_t)ndir);
  /* SH2A (fmls.cpp:175-176), accumulated by fused multiply-add in ascending coefficient order: the
   * reference's Eigen product rounds differently in the last bit, which can only reorder directions
   * of equal amplitude to within that bit */
  for (int d = 0; d < ndir; d++) {
    double s = 0.0;
    for (int k = 0; k < nsh; k++) s = fma(SHT[(size_t)d * nsh + k], sh_in[k], s);
    values[d] = s;
    order[d].v = s;
    order[d].i = d;
  }
  qsort(order, (size_t)ndir, sizeof(vo_pair), cmp_absdesc);
  int nl = 0;
  if (order[0].v <= 0.0) {
    free(values); free(order);
    if (lut) for (int d = 0; d < ndir; d++) lut[d] = 0;
    return 0;
  }
  lobe* lobes = (lobe*)calloc(FMLS_MAX_LOBES, sizeof(lobe));
  int* owner = (int*)malloc(sizeof(int) * (size_t)ndir);   /* lobe index whose mask holds d, -1 */
  for (int d = 0; d < ndir; d++) owner[d] = -1;
  int* retro_d = (int*)malloc(sizeof(int) * (size_t)ndir);
  int* retro_l = (int*)malloc(sizeof(int) * (size_t)ndir);
  int nretro = 0;
  int overflow = 0;
  for (int oi = 0; oi < ndir && !overflow; oi++) {
    const double v = order[oi].v;
    const int d = order[oi].i;
    int adjl[FMLS_MAX_LOBES];
    int na = 0;
    for (int l = 0; l < nl; l++) {
      if (!(((v <= 0.0) && lobes[l].neg) || ((v > 0.0) && !lobes[l].neg))) continue;
      int is_adj = 0;
      for (int a = 0; a < FMLS_MAX_ADJ && !is_adj; a++) {
        const int j = adj[d * FMLS_MAX_ADJ + a];
        if (j >= 0 && owner[j] == l) is_adj = 1;
      }
      if (is_adj) adjl[na++] = l;
    }
    if (na == 0) {
      if (nl == FMLS_MAX_LOBES) { overflow = 1; break; }
      lobe* L = &lobes[nl];
      memset(L, 0, sizeof(lobe));
      L->max_peak = fabs(v);
      L->npeaks = 1;
      memcpy(L->peaks[0], unit + 3 * d, sizeof(double) * 3);
      for (int c = 0; c < 3; c++) L->mean[c] = unit[3 * d + c] * fabs(v) * weights[d];
      L->integral = fabs(v * weights[d]);
      L->neg = v <= 0.0;
      owner[d] = nl;
      nl++;
    } else if (na == 1) {
      owner[d] = adjl[0];
      lobe_add(&lobes[adjl[0]], unit, d, v, weights[d]);
    } else {
      if (fabs(v) / lobes[adjl[na - 1]].max_peak > merge_ratio) {
        /* adjl is in ascending order already; merge all into the first (fmls.h:112-127) */
        lobe* A = &lobes[adjl[0]];
        for (int k = 1; k < na; k++) {
          lobe* B = &lobes[adjl[k]];
          for (int dd = 0; dd < ndir; dd++) if (owner[dd] == adjl[k]) owner[dd] = adjl[0];
          if (B->max_peak > A->max_peak) {
            A->max_peak = B->max_peak;
            double tmp[FMLS_MAX_PEAKS][3];
            int np = 0;
            for (int p = 0; p < B->npeaks && np < FMLS_MAX_PEAKS; p++) memcpy(tmp[np++], B->peaks[p], sizeof(double) * 3);
            for (int p = 0; p < A->npeaks && np < FMLS_MAX_PEAKS; p++) memcpy(tmp[np++], A->peaks[p], sizeof(double) * 3);
            memcpy(A->peaks, tmp, sizeof(tmp));
            A->npeaks = np;
          } else {
            for (int p = 0; p < B->npeaks && A->npeaks < FMLS_MAX_PEAKS; p++) memcpy(A->peaks[A->npeaks++], B->peaks[p], sizeof(double) * 3);
          }
          const double mult = (A->mean[0] * B->mean[0] + A->mean[1] * B->mean[1] + A->mean[2] * B->mean[2]) > 0.0 ? 1.0 : -1.0;
          for (int c = 0; c < 3; c++) A->mean[c] += B->mean[c] * B->integral * mult;
          A->integral += B->integral;
        }
        owner[d] = adjl[0];
        lobe_add(A, unit, d, v, weights[d]);
        for (int r = 0; r < nretro; r++) {
          int modified = 0;
          for (int k = 1; k < na; k++)
            if (retro_l[r] == adjl[k]) { retro_l[r] = adjl[0]; modified = 1; }
          if (!modified) {
            int li = retro_l[r];
            for (int k = na - 1; k; k--)
              if (adjl[k] < li) --li;
            retro_l[r] = li;
          }
        }
        /* erase the merged lobes (highest first) and renumber the owners */
        for (int k = na - 1; k; k--) {
          const int e = adjl[k];
          for (int l = e; l < nl - 1; l++) lobes[l] = lobes[l + 1];
          nl--;
          for (int dd = 0; dd < ndir; dd++)
            if (owner[dd] > e) owner[dd]--;
        }
      } else {
        retro_d[nretro] = d;
        retro_l[nretro] = adjl[0];
        nretro++;
      }
    }
  }
  if (overflow) {
    free(values); free(order); free(lobes); free(owner); free(retro_d); free(retro_l);
    return -1;
  }
  for (int r = 0; r < nretro; r++) {
    owner[retro_d[r]] = retro_l[r];
    lobe_add(&lobes[retro_l[r]], unit, retro_d[r], values[retro_d[r]], weights[retro_d[r]]);
  }
  /* negative and small lobes dropped, peaks revised, thresholds (fmls.cpp:256-301) */
  int nk = 0;
  for (int l = 0; l < nl; l++) {
    lobe* L = &lobes[l];
    if (L->neg || L->integral < integral_threshold) {
      for (int dd = 0; dd < ndir; dd++) if (owner[dd] == l) owner[dd] = -2;
      continue;
    }
    for (int p = 0; p < L->npeaks && al_table; p++) {   /* no table: peaks left as found */
      double nd[3] = {L->peaks[p][0], L->peaks[p][1], L->peaks[p][2]};
      const double nv = orc_fmls_get_peak(sh_in, lmax, nd, al_table, al_ndir);
      if (isfinite(nv) && isfinite(nd[0]) && isfinite(nd[1]) && isfinite(nd[2])) {
        double max_dp = 0.0;
        int nearest = L->npeaks;
        for (int j = 0; j < L->npeaks; j++) {
          const double dp = fabs(nd[0] * L->peaks[j][0] + nd[1] * L->peaks[j][1] + nd[2] * L->peaks[j][2]);
          if (dp > max_dp) { max_dp = dp; nearest = j; }
        }
        if (nearest == p) {
          const int closest = orc_fmls_select(look, unit, nd);
          if (owner[closest] == l) {
            memcpy(L->peaks[p], nd, sizeof(nd));
            if (!p) L->max_peak = nv;
          }
        }
      }
    }
    if (L->max_peak < peak_value_threshold) {
      for (int dd = 0; dd < ndir; dd++) if (owner[dd] == l) owner[dd] = -2;
      continue;
    }
    const double nrm = sqrt(L->mean[0] * L->mean[0] + L->mean[1] * L->mean[1] + L->mean[2] * L->mean[2]);
    if (nk < max_out) {
      double* o = out + (size_t)nk * 8;
      for (int c = 0; c < 3; c++) o[c] = L->mean[c] / nrm;
      for (int c = 0; c < 3; c++) o[3 + c] = L->peaks[0][c];
      o[6] = L->integral;
      o[7] = L->max_peak;
    }
    for (int dd = 0; dd < ndir; dd++) if (owner[dd] == l) owner[dd] = 1000000 + nk;
    nk++;
  }
  if (lut)
    for (int d = 0; d < ndir; d++) lut[d] = owner[d] >= 1000000 ? owner[d] - 1000000 : nk;
  free(values); free(order); free(lobes); free(owner); free(retro_d); free(retro_l);
  return nk;
}
