// FOD segmentation on gfx950: fod2fixel's FMLS (src/dwi/fmls.cpp:166-377) for a batch of voxels.
//
// Per chunk of voxels:
//   * fmls_amp_kernel: the FOD amplitudes on the segmentation directions (1281 for fod2fixel) as
//     one FP64 GEMM, amps[v][d] = sum_k SHT[d][k] F_v[k] accumulated by FMA in ascending k,
//     LDS-tiled (128 directions x 64 voxels per workgroup, 8 x 4 outputs per thread), one row per
//     voxel;
//   * fmls_sort_kernel: a wavefront per voxel checks that the largest |amplitude| is positive and
//     sorts the positive directions by decreasing amplitude, the lowest index first among equal
//     amplitudes (the reference's multimap order), by a bitonic sort in registers;
//   * fmls_grow_kernel<Lab8>: lobe growth, one lane per voxel, step for step as the reference: a
//     direction adjacent to no positive lobe seeds one, adjacent to one joins it, adjacent to
//     several either merges them (bridge amplitude over the peak of the highest-numbered adjacent
//     lobe above the merge ratio) or waits, outside every lobe, for its assignment to the
//     lowest-numbered of them after the loop.  Each lane's lobe labels are an 8-bit column in LDS
//     (as many voxels per workgroup as let two share a CU: 63 in 160 KB); the sorted lists and
//     direction tables are loaded a block of steps at a time; lobe sums are kept in a register
//     cache.  A voxel of more than 127 lobes
//     is left to fmls_grow_kernel<Lab16> (16-bit labels, 31 voxels per workgroup), which grows it
//     again from the unchanged sorted lists.  The final labels go to the work arrays;
//   * fmls_block_sums, fmls_scan_sums, fmls_lobe_items: the chunk's lobes listed as work items;
//   * fmls_revise_kernel<LMAX>: one lane per lobe, the lanes striding over the items: the
//     integral threshold, the Newton peak search from each of its peaks (Math::SH::get_peak in
//     double with the precomputed Legendre table, SH.h:398-559), a revision kept only when it stays
//     nearest to that peak and inside the lobe (nearest direction through the FastLookupSet grid),
//     the peak threshold, the fixel direction;
//   * fmls_finalize_kernel: each voxel's fixel count and summary; the kept lobes stay in the work
//     arrays;
//   * fmls_block_sums, fmls_scan_sums: the fixel offset of each block of 64 voxels, continuing
//     the running total of the earlier chunks;
//   * fmls_emit: each voxel's offset in voxel order (the order fod2fixel's index image uses,
//     cmd/fod2fixel.cpp:239-280; 0 for voxels without fixels) and its fixels.
// The work arrays hold every table a voxel can need (as many lobes and peaks as directions), so
// no voxel is deferred.  The growth and revision have no cross-lane operation (their LDS is a
// column per lane): the cooperation between lanes is the sort's exchanges within a wavefront and
// the workgroup scans of the offset kernels.
// Negative directions only ever form negative lobes, which the reference discards and which take
// no part in positive lobes' adjacency, order, merges or indices, so they are not visited.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fmls.h"

// LDS of the growth kernels, sized at launch: `lanes` columns of ndir labels
#ifndef FMLS_GROW_LDS_ARRAY
extern __shared__ unsigned char fmls_grow_lds[];
#define FMLS_GROW_LDS_ARRAY fmls_grow_lds
#endif

namespace mrcsd {
namespace {

// One voxel's view of a lane-interleaved work array: element i at p[i * s + o], with p and s the
// same for every lane (scalar registers) and o the voxel's slot.  Indices stay below 2^31: at most
// 3 x kFmlsMaxDirs elements of a chunk of at most 2^17 slots.
template <class T>
struct Lane {
  T* p;
  uint32_t s, o;
  __device__ __forceinline__ T& operator[](uint32_t i) const { return p[i * s + o]; }
};

__device__ __forceinline__ int fm_mpos(int l, int m) { return l * l / 4 + m; }
__device__ __forceinline__ int fm_sh(int l, int m) { return l * (l + 1) / 2 + m; }

// Math::SH::derivatives with PrecomputedAL<double> (SH.h:568-647): out = {amplitude, d/del,
// d/daz, d2/del2, d2/deldaz, d2/daz2}.  sh: the voxel's float32 coefficients; el the elevation,
// (caz, saz) the cosine and sine of the azimuth, sel the sine of the elevation.  The Legendre
// values of the elevation, interpolated between two rows of the precomputed table
// (PrecomputedAL::set / get, SH.h:436-470), go to this lane's own column of an LDS table (entry k
// at AL[kFmlsBlock k]), read by no other lane.  cos(m az), sin(m az) come from the m = 1 values by
// angle addition; up to lmax 8 the loops over l and m are unrolled, so their square-root
// constants fold.
template <int LMAX>
__device__ void fmls_derivatives(const float* sh, double el, double caz, double saz, double sel,
                                 const FmlsDeviceSet& S, double* AL, double* out) {
  constexpr int NAL = (LMAX / 2 + 1) * (LMAX / 2 + 1);
  const bool atpole = sel < 1e-4;
  {
    const double inc = M_PI / (S.al_ndir - 1);
    double f2 = el / inc, f1;
    int i;
    if (!(f2 >= 0.0)) {             // (also a NaN elevation: the first row, never out of bounds)
      i = 0;
      f1 = 1.0;
      f2 = 0.0;
    } else if (f2 >= (double)(S.al_ndir - 1)) {
      i = S.al_ndir - 1;
      f1 = 1.0;
      f2 = 0.0;
    } else {
      i = (int)f2;
      f2 -= i;
      f1 = 1.0 - f2;
    }
    const double* p1 = S.al + (size_t)i * NAL;
    const double* p2 = p1 + (f2 != 0.0 ? NAL : 0);
    for (int k = 0; k < NAL; ++k) {
      double v = f1 * p1[k];
      if (f2 != 0.0) v += f2 * p2[k];
      AL[kFmlsBlock * k] = v;
    }
  }
#define FM_AL(k) AL[kFmlsBlock * (k)]
  double amplitude = (double)sh[0] * FM_AL(0);
  double dSH_del = 0.0, dSH_daz = 0.0, d2SH_del2 = 0.0, d2SH_deldaz = 0.0, d2SH_daz2 = 0.0;
#pragma unroll
  for (int l = 2; l <= LMAX; l += 2) {
    const double v = (double)sh[fm_sh(l, 0)];
    amplitude += v * FM_AL(fm_mpos(l, 0));
    dSH_del += v * sqrt((double)(l * (l + 1))) * FM_AL(fm_mpos(l, 1));
    d2SH_del2 += v * (sqrt((double)(l * (l + 1) * (l - 1) * (l + 2))) * FM_AL(fm_mpos(l, 2)) -
                      l * (l + 1) * FM_AL(fm_mpos(l, 0))) / 2.0;
  }
  double cm = caz, sm = saz;   // cos(m az), sin(m az)
  constexpr int kUnrollM = LMAX <= 8 ? LMAX : 1;   // above lmax 8 the unrolled loop spills
#pragma unroll kUnrollM
  for (int m = 1; m <= LMAX; ++m) {
    const double caz_m = M_SQRT2 * cm;
    const double saz_m = M_SQRT2 * sm;
#pragma unroll
    for (int l = (m & 1) ? m + 1 : m; l <= LMAX; l += 2) {
      const double vp = (double)sh[fm_sh(l, m)];
      const double vm = (double)sh[fm_sh(l, -m)];
      const double alm = FM_AL(fm_mpos(l, m));
      amplitude += (vp * caz_m + vm * saz_m) * alm;
      double tmp = sqrt((double)((l + m) * (l - m + 1))) * FM_AL(fm_mpos(l, m - 1));
      if (l > m) tmp -= sqrt((double)((l - m) * (l + m + 1))) * FM_AL(fm_mpos(l, m + 1));
      tmp /= -2.0;
      dSH_del += (vp * caz_m + vm * saz_m) * tmp;
      double tmp2 = -((l + m) * (l - m + 1) + (l - m) * (l + m + 1)) * alm;
      if (m == 1) tmp2 -= sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * FM_AL(fm_mpos(l, 1));
      else tmp2 += sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * FM_AL(fm_mpos(l, m - 2));
      if (l > m + 1) tmp2 += sqrt((double)((l - m) * (l + m + 1) * (l - m - 1) * (l + m + 2))) * FM_AL(fm_mpos(l, m + 2));
      tmp2 /= 4.0;
      d2SH_del2 += (vp * caz_m + vm * saz_m) * tmp2;
      if (atpole) {
        dSH_daz += (vm * caz_m - vp * saz_m) * tmp;
      } else {
        d2SH_deldaz += m * (vm * caz_m - vp * saz_m) * tmp;
        dSH_daz += m * (vm * caz_m - vp * saz_m) * alm;
        d2SH_daz2 -= (vp * caz_m + vm * saz_m) * m * m * alm;
      }
    }
    const double cn = cm * caz - sm * saz;
    sm = sm * caz + cm * saz;
    cm = cn;
  }
  if (!atpole) {
    dSH_daz /= sel;
    d2SH_deldaz /= sel;
    d2SH_daz2 /= sel * sel;
  }
#undef FM_AL
  out[0] = amplitude; out[1] = dSH_del; out[2] = dSH_daz;
  out[3] = d2SH_del2; out[4] = d2SH_deldaz; out[5] = d2SH_daz2;
}

// Math::SH::get_peak (SH.h:514-559): Newton steps along the gradient, at most 0.2 rad each,
// until a step is below 1e-4 rad; NaN after 50 steps.  The azimuth's and elevation's cosines and
// sines come from the unit direction's components (x / rho, y / rho, z, rho with rho = hypot(x, y));
// only the elevation itself, which indexes the Legendre table, takes an atan2.  At the pole
// (rho = 0) the azimuth is atan2(y, x) as the reference takes it.
template <int LMAX>
__device__ double fmls_get_peak(const float* sh, double* dir, const FmlsDeviceSet& S, double* AL) {
  const double MAX_DIR_CHANGE = 0.2, ANGLE_TOLERANCE = 1e-4;
  for (int it = 0; it < 50; ++it) {
    const double rho = hypot(dir[0], dir[1]);
    const double el = atan2(rho, dir[2]);
    double ca, sa;
    if (rho > 0.0) {
      ca = dir[0] / rho;
      sa = dir[1] / rho;
    } else {
      const double az = atan2(dir[1], dir[0]);
      ca = cos(az);
      sa = sin(az);
    }
    const double r = hypot(rho, dir[2]);
    const double ce = dir[2] / r, se = rho / r;
    double d[6];
    fmls_derivatives<LMAX>(sh, el, ca, sa, se, S, AL, d);
    double del = sqrt(d[1] * d[1] + d[2] * d[2]);
    double daz = 0.0;
    if (del != 0.0) {
      daz = d[2] / del;
      del = d[1] / del;
    }
    const double dSH_dt = daz * d[2] + del * d[1];
    const double d2SH_dt2 = daz * daz * d[5] + 2.0 * daz * del * d[4] + del * del * d[3];
    double dt = d2SH_dt2 != 0.0 ? (-dSH_dt / d2SH_dt2) : MAX_DIR_CHANGE;
    if (dt < 0.0) dt = -dt;
    if (dt > MAX_DIR_CHANGE) dt = MAX_DIR_CHANGE;
    del *= dt;
    daz *= dt;
    dir[0] += del * ca * ce - daz * sa;
    dir[1] += del * sa * ce + daz * ca;
    dir[2] -= del * se;
    const double nrm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] /= nrm;
    dir[1] /= nrm;
    dir[2] /= nrm;
    if (dt < ANGLE_TOLERANCE) return d[0];
  }
  dir[0] = dir[1] = dir[2] = NAN;
  return NAN;
}

// FastLookupSet::select_direction (directions/set.cpp:336-355): the nearest direction among the
// candidates of the (azimuth, elevation) grid cell
__device__ int fmls_select(const FmlsDeviceSet& S, const double* p) {
  const double az = atan2(p[1], p[0]);
  const double el = acos(p[2]);
  // grid cell; indices clamped to the grid (an elevation of NaN from |z| rounded above 1 included)
  double fa = floor((az - S.az_begin) / S.az_step), fe = floor((el - S.el_begin) / S.el_step);
  fa = fa >= 0.0 ? (fa <= (double)(S.n_az - 1) ? fa : (double)(S.n_az - 1)) : 0.0;
  fe = fe >= 0.0 ? (fe <= (double)(S.n_el - 1) ? fe : (double)(S.n_el - 1)) : 0.0;
  const size_t g = (size_t)fa * (size_t)S.n_el + (size_t)fe;
  const int b = S.cell_start[g], e = S.cell_start[g + 1];
  int best = S.cell_dirs[b];
  double mx = fabs(p[0] * S.unit[3 * best] + p[1] * S.unit[3 * best + 1] + p[2] * S.unit[3 * best + 2]);
  for (int k = b + 1; k < e; ++k) {
    const int d = S.cell_dirs[k];
    const double v = fabs(p[0] * S.unit[3 * d] + p[1] * S.unit[3 * d + 1] + p[2] * S.unit[3 * d + 2]);
    if (v > mx) {
      mx = v;
      best = d;
    }
  }
  return best;
}

// One voxel's work arrays.  Lobe records follow FOD_lobe (fmls.h:61-137); a lobe's peaks are the
// seeds of the lobes merged into it, kept as a linked list through nxt[seed direction] in the
// reference's order.
struct Tabs {
  Lane<int16_t> lab;
  Lane<uint32_t> ordm;
  Lane<int16_t> nxt;
  Lane<double> peak, mean, integ;
  Lane<int16_t> head, tail, np;
  Lane<double> rev;
};

__device__ Tabs tabs_of(const FmlsWork& W, size_t slot, size_t cs) {
  const uint32_t o = (uint32_t)slot, c = (uint32_t)cs;
  return Tabs{{W.lab, c, o},  {W.ordm, c, o},  {W.nxt, c, o}, {W.peak, c, o}, {W.mean, c, o},
              {W.integ, c, o}, {W.head, c, o}, {W.tail, c, o}, {W.np, c, o},  {W.rev, c, o}};
}

// a direction's adjacency row (8 int16 entries), loaded as one 16-byte access
struct alignas(16) AdjRow {
  uint32_t w[4];
};
__device__ __forceinline__ AdjRow adj_row(const FmlsDeviceSet& S, int d) {
  return *reinterpret_cast<const AdjRow*>(S.adj + d * 8);
}

// What a growth step needs of its direction d from the direction tables: unit vector,
// integration weight, adjacency row.
struct DirInfo {
  double u0, u1, u2, w;
  AdjRow adj;
};
__device__ __forceinline__ DirInfo dir_info(const FmlsDeviceSet& S, int d) {
  const double* u = S.unit + 3 * d;
  return DirInfo{u[0], u[1], u[2], S.weights[d], adj_row(S, d)};
}

// Write-back cache of lobe records in registers: the running direction sum and integral of up to
// four lobes, entry 0 the lobe last touched (a voxel mostly grows the lobe it grew last, which then
// costs no select or memory access), entries 1..3 the ones touched before it, most recent first.
// Entries leave through entry 3, written back to the work arrays.  tag -1: empty.
struct LobeCache {
  int tag[4];
  double m0[4], m1[4], m2[4], integ[4];
};

__device__ __forceinline__ void lc_init(LobeCache& c) {
#pragma unroll
  for (int q = 0; q < 4; ++q) c.tag[q] = -1;
}

__device__ __forceinline__ void lc_store(const Tabs& T, int l, double m0, double m1, double m2, double integ) {
  T.mean[3 * l] = m0;
  T.mean[3 * l + 1] = m1;
  T.mean[3 * l + 2] = m2;
  T.integ[l] = integ;
}

// every entry back to the work arrays, the cache emptied
__device__ void lc_flush(const Tabs& T, LobeCache& c) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (c.tag[q] >= 0) lc_store(T, c.tag[q], c.m0[q], c.m1[q], c.m2[q], c.integ[q]);
    c.tag[q] = -1;
  }
}

// lobe l to entry 0: found among entries 1..3, or read from the work arrays (unless `fresh`, a
// new lobe whose values the caller sets) after entry 3 is written back; the entries before it
// move down one
__device__ __forceinline__ void lc_take(const Tabs& T, LobeCache& c, int l, bool fresh) {
  if (c.tag[0] == l) return;
  int q = 3;   // the entry l comes from (3: none, evicted)
#pragma unroll
  for (int k = 3; k >= 1; --k)
    if (c.tag[k] == l) q = k;
  double n0 = 0.0, n1 = 0.0, n2 = 0.0, ni = 0.0;
  int tq = -1;
#pragma unroll
  for (int k = 1; k < 4; ++k)
    if (k == q) {
      tq = c.tag[k];
      n0 = c.m0[k];
      n1 = c.m1[k];
      n2 = c.m2[k];
      ni = c.integ[k];
    }
  if (tq != l) {
    if (c.tag[3] >= 0) lc_store(T, c.tag[3], c.m0[3], c.m1[3], c.m2[3], c.integ[3]);
    if (!fresh) {
      n0 = T.mean[3 * l];
      n1 = T.mean[3 * l + 1];
      n2 = T.mean[3 * l + 2];
      ni = T.integ[l];
    }
  }
#pragma unroll
  for (int k = 3; k >= 1; --k)
    if (k <= q) {
      c.tag[k] = c.tag[k - 1];
      c.m0[k] = c.m0[k - 1];
      c.m1[k] = c.m1[k - 1];
      c.m2[k] = c.m2[k - 1];
      c.integ[k] = c.integ[k - 1];
    }
  c.tag[0] = l;
  c.m0[0] = n0;
  c.m1[0] = n1;
  c.m2[0] = n2;
  c.integ[0] = ni;
}

// a new lobe seeded by direction d of amplitude v (fmls.cpp:199-201, FOD_lobe's constructor
// fmls.h:64-76)
__device__ void lobe_new(const Tabs& T, int l, int d, double v, const DirInfo& di, LobeCache& c) {
  lc_take(T, c, l, true);
  c.m0[0] = di.u0 * fabs(v) * di.w;
  c.m1[0] = di.u1 * fabs(v) * di.w;
  c.m2[0] = di.u2 * fabs(v) * di.w;
  c.integ[0] = fabs(v * di.w);
  T.peak[l] = fabs(v);
  T.head[l] = (int16_t)d;
  T.tail[l] = (int16_t)d;
  T.np[l] = 1;
  T.nxt[d] = -1;
}

// FOD_lobe::add (fmls.h:88-97), through the cache
__device__ void lobe_add(const Tabs& T, int l, double v, const DirInfo& di, LobeCache& c) {
  lc_take(T, c, l, false);
  const double m0 = c.m0[0], m1 = c.m1[0], m2 = c.m2[0];
  const double mult = (m0 * di.u0 + m1 * di.u1 + m2 * di.u2) > 0.0 ? 1.0 : -1.0;
  const double av = fabs(v);
  c.m0[0] = m0 + di.u0 * mult * av * di.w;
  c.m1[0] = m1 + di.u1 * mult * av * di.w;
  c.m2[0] = m2 + di.u2 * mult * av * di.w;
  c.integ[0] = c.integ[0] + fabs(v * di.w);
}

// FOD_lobe::merge (fmls.h:112-127): B into A
__device__ void lobe_merge(const Tabs& T, int A, int B) {
  if (T.peak[B] > T.peak[A]) {
    T.peak[A] = T.peak[B];
    T.nxt[T.tail[B]] = T.head[A];    // B's peaks first
    T.head[A] = T.head[B];
  } else {
    T.nxt[T.tail[A]] = T.head[B];    // B's peaks after A's
    T.tail[A] = T.tail[B];
  }
  T.np[A] = (int16_t)(T.np[A] + T.np[B]);
  const double a0 = T.mean[3 * A], a1 = T.mean[3 * A + 1], a2 = T.mean[3 * A + 2];
  const double b0 = T.mean[3 * B], b1 = T.mean[3 * B + 1], b2 = T.mean[3 * B + 2];
  const double mult = (a0 * b0 + a1 * b1 + a2 * b2) > 0.0 ? 1.0 : -1.0;
  const double ib = T.integ[B];
  T.mean[3 * A] = a0 + b0 * ib * mult;
  T.mean[3 * A + 1] = a1 + b1 * ib * mult;
  T.mean[3 * A + 2] = a2 + b2 * ib * mult;
  T.integ[A] = T.integ[A] + ib;
}

__device__ void lobe_copy(const Tabs& T, int dst, int src) {
  T.peak[dst] = T.peak[src];
  T.mean[3 * dst] = T.mean[3 * src];
  T.mean[3 * dst + 1] = T.mean[3 * src + 1];
  T.mean[3 * dst + 2] = T.mean[3 * src + 2];
  T.integ[dst] = T.integ[src];
  T.head[dst] = T.head[src];
  T.tail[dst] = T.tail[src];
  T.np[dst] = T.np[src];
}

// A lane's column of lobe labels in the growth kernels' LDS: entry d at element o + d * s (32-bit
// indices, so addresses stay LDS offsets).  -1: no lobe (not visited, or negative); l >= 0: lobe
// l; -2 - l: waiting for lobe l.  CAP: the lobes the type can number.
template <class V, int CAP>
struct LdsLab {
  uint32_t o, s;
  static constexpr int cap = CAP;
  __device__ __forceinline__ int get(int d) const {
    return reinterpret_cast<V*>(FMLS_GROW_LDS_ARRAY)[o + (uint32_t)d * s];
  }
  __device__ __forceinline__ void set(int d, int v) const {
    reinterpret_cast<V*>(FMLS_GROW_LDS_ARRAY)[o + (uint32_t)d * s] = (V)v;
  }
};
using Lab8 = LdsLab<int8_t, 127>;        // waiting labels down to -128: lobes 0..126
using Lab16 = LdsLab<int16_t, 32767>;    // any direction count

// The ascending distinct lobes among d's neighbour labels (INT_MAX pads), each lobe merged into
// the lowest-numbered one in ascending order, d added to it, and the labels and lobe records
// renumbered over the erased records (fmls.cpp:203-237); only lobe merge ratios below 1 can
// merge.  The caller's cache is flushed; this uses its own.  Returns the new lobe count.
template <class Lab>
__device__ __forceinline__ int grow_merge(const FmlsDeviceSet& S, const Tabs& T, Lab lab, const int* lbs, int d, double v,
                                       double u0, double u1, double u2, double w, int nl) {
  const DirInfo di{u0, u1, u2, w, AdjRow{}};
  LobeCache lc;
  lc_init(lc);
  int al[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) al[q] = INT_MAX;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int lb = lbs[r];
    if (lb < 0) continue;
    bool seen = false;
#pragma unroll
    for (int q = 0; q < 8; ++q) seen = seen || al[q] == lb;
    if (seen) continue;
#pragma unroll
    for (int q = 7; q > 0; --q) {
      if (al[q - 1] > lb) al[q] = al[q - 1];
      else if (al[q] > lb) al[q] = lb;
    }
    if (al[0] > lb) al[0] = lb;
  }
  const int A = al[0];
#pragma unroll
  for (int q = 1; q < 8; ++q)
    if (al[q] != INT_MAX) lobe_merge(T, A, al[q]);
  lab.set(d, A);
  lobe_add(T, A, v, di, lc);
  lc_flush(T, lc);   // records move below
  // merged lobes become A, the lobes above them move down over the erased records; the same for
  // the lobes waiting directions are assigned to
  for (int dd = 0; dd < S.ndir; ++dd) {
    const int lb = lab.get(dd);
    const bool wait = lb <= -2;
    int l = wait ? -2 - lb : lb;
    if (l < 0) continue;
    bool merged = false;
    int sh = 0;
#pragma unroll
    for (int q = 1; q < 8; ++q)
      if (al[q] != INT_MAX) {
        merged = merged || l == al[q];
        sh += al[q] < l ? 1 : 0;
      }
    if (!merged && sh == 0) continue;
    l = merged ? A : l - sh;
    lab.set(dd, wait ? -2 - l : l);
  }
#pragma unroll
  for (int q = 7; q >= 1; --q) {
    if (al[q] == INT_MAX) continue;
    for (int l = al[q]; l < nl - 1; ++l) lobe_copy(T, l, l + 1);
    --nl;
  }
  return nl;
}

// Lobe growth and the retrospective assignments of one voxel (fmls.cpp:188-252) from its P > 0
// positive directions in visiting order (ordm, with their visited-before neighbour masks, and
// their amplitudes in rev), its labels in `lab`; returns the number of positive lobes, or -1 when
// the voxel needs more lobes than `lab` can number (nothing in ordm or rev is changed, so the
// voxel can be grown again with wider labels).
//
// The global loads of a block of eight steps (visiting order, neighbour masks, amplitudes, then
// the directions' unit vectors, weights and adjacency rows) are issued before its first step.
// Labels live in LDS and lobe sums in the register cache, so a step usually issues no global
// access of its own.
template <class Lab>
__device__ int fmls_grow_voxel(const FmlsDeviceSet& S, const Tabs& T, const Lab& lab, int P) {
  constexpr int KA = 8;
  for (int d = 0; d < S.ndir; ++d) lab.set(d, -1);
  int nl = 0;
  LobeCache lc;
  lc_init(lc);
  const bool can_merge = S.merge_ratio < 1.0;   // (a direction's amplitude never exceeds a visited peak)
  // Blocks of KA steps: the block's visiting order, masks and amplitudes are loaded first, then
  // its directions' tables, all before its first step, and consumed in straight-line code -- no
  // load is in flight across the loop's back edge (the memory counter is in order, so a load
  // carried over would make the compiler wait for everything at the loop head).  Indices past the
  // end are clamped to the last entry: in bounds, never used.
  for (int i0 = 0; i0 < P; i0 += KA) {
    int ra[KA];      // direction | visited-before neighbour mask << 16
    double rv[KA];   // amplitude
    DirInfo rb[KA];
#pragma unroll
    for (int j = 0; j < KA; ++j) {
      const int ia = i0 + j < P ? i0 + j : P - 1;
      ra[j] = (int)T.ordm[ia];
      rv[j] = T.rev[ia];
    }
#pragma unroll
    for (int j = 0; j < KA; ++j) rb[j] = dir_info(S, ra[j] & 0xffff);
#pragma unroll
    for (int j = 0; j < KA; ++j) {
      const int i = i0 + j;
      if (i < P) {
        const int d = ra[j] & 0xffff;
        const unsigned seen_before = (unsigned)ra[j] >> 16;
        const double v = rv[j];
        const DirInfo& di = rb[j];
        // the lowest- and highest-numbered lobes among d's neighbours; only neighbours visited
        // before d can carry one (the others still hold -1), so only their labels are read, and
        // waiting directions (<= -2) belong to no lobe yet
        // (every label is read -- padding entries read entry 0 -- and the mask picks; the
        // highest label by signed max, where -1 and waiting labels stay below every lobe; the
        // lowest by unsigned min, where they stay above)
        int lbs[8];
        int mx = -1;
        unsigned umn = 0xffffffffu;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int jr = (int16_t)((di.adj.w[r / 2] >> (16 * (r % 2))) & 0xffffu);
          const int raw = lab.get(jr < 0 ? 0 : jr);
          lbs[r] = (seen_before >> r) & 1u ? raw : -1;
          mx = max(mx, lbs[r]);
          umn = min(umn, (unsigned)lbs[r]);
        }
        const int mn = (int)umn;
        if (mx < 0) {
          if (nl == Lab::cap) return -1;
          lobe_new(T, nl, d, v, di, lc);
          lab.set(d, nl);
          ++nl;
        } else if (mn == mx) {
          lab.set(d, mn);
          lobe_add(T, mn, v, di, lc);
        } else if (can_merge && fabs(v) / T.peak[mx] > S.merge_ratio) {
          lc_flush(T, lc);   // the merges read and rewrite records
          int lbc[8];
#pragma unroll
          for (int r = 0; r < 8; ++r) lbc[r] = lbs[r];
          nl = grow_merge(S, T, lab, lbc, d, v, di.u0, di.u1, di.u2, di.w, nl);
        } else {
          lab.set(d, -2 - mn);
        }
      }
    }
  }

  // retrospective assignments in visiting order (fmls.cpp:251-252): the waiting directions, whose
  // labels still name the lobe they wait for (renumbered by later merges, never assigned)
  for (int i = 0; i < P; ++i) {
    const int d = (int)(T.ordm[i] & 0xffffu);
    const int lb = lab.get(d);
    if (lb <= -2) {
      const int l = -2 - lb;
      lab.set(d, l);
      lobe_add(T, l, T.rev[i], dir_info(S, d), lc);
    }
  }
  lc_flush(T, lc);
  return nl;
}

// The per-lobe steps of the segmentation (fmls.cpp:254-292) for lobe l of a voxel, the final labels
// in the work arrays: the integral threshold, the Newton search from each of its peaks in the
// reference's order, the nearest-peak and inside-the-lobe checks, the peak threshold, the fixel
// direction; np = -1 when a threshold drops it.  Peak p, the p-th seed s_p on the lobe's list,
// keeps its direction at rev[3 s_p]: seeds are distinct directions, so the lobes of a voxel use
// disjoint entries and can be revised independently.
template <int LMAX>
__device__ void fmls_revise_lobe(const FmlsDeviceSet& S, const Tabs& T, int l, const float* f, double* AL) {
  const int np = T.np[l];
  const int head = T.head[l];
  for (int p = 0, s = head; p < np && s >= 0; ++p, s = T.nxt[s]) {
    T.rev[3 * s] = S.unit[3 * s];
    T.rev[3 * s + 1] = S.unit[3 * s + 1];
    T.rev[3 * s + 2] = S.unit[3 * s + 2];
  }
  bool keep = !(T.integ[l] < S.integral_threshold);
  double maxp = T.peak[l];
  if (keep) {
    for (int p = 0, sp = head; p < np && sp >= 0; ++p, sp = T.nxt[sp]) {
      double nd[3] = {T.rev[3 * sp], T.rev[3 * sp + 1], T.rev[3 * sp + 2]};
      const double nvl = fmls_get_peak<LMAX>(f, nd, S, AL);
      if (isfinite(nvl) && isfinite(nd[0]) && isfinite(nd[1]) && isfinite(nd[2])) {
        double max_dp = 0.0;
        int nearest = np;
        for (int q = 0, sq = head; q < np && sq >= 0; ++q, sq = T.nxt[sq]) {
          const double dp = fabs(nd[0] * T.rev[3 * sq] + nd[1] * T.rev[3 * sq + 1] + nd[2] * T.rev[3 * sq + 2]);
          if (dp > max_dp) {
            max_dp = dp;
            nearest = q;
          }
        }
        if (nearest == p && T.lab[fmls_select(S, nd)] == l) {
          T.rev[3 * sp] = nd[0];
          T.rev[3 * sp + 1] = nd[1];
          T.rev[3 * sp + 2] = nd[2];
          if (p == 0) maxp = nvl;
        }
      }
    }
    keep = !(maxp < S.peak_threshold);
  }
  if (keep) {
    if (S.dir_from_peak) {
      T.mean[3 * l] = T.rev[3 * head];
      T.mean[3 * l + 1] = T.rev[3 * head + 1];
      T.mean[3 * l + 2] = T.rev[3 * head + 2];
    } else {
      const double m0 = T.mean[3 * l], m1 = T.mean[3 * l + 1], m2 = T.mean[3 * l + 2];
      const double nrm = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
      T.mean[3 * l] = m0 / nrm;
      T.mean[3 * l + 1] = m1 / nrm;
      T.mean[3 * l + 2] = m2 / nrm;
    }
    T.peak[l] = maxp;
  } else {
    T.np[l] = -1;
  }
}

// The chunk's lobes as work items, voxel by voxel: items[o + l] = slot << 11 | l for the voxel's
// lobes l, o its first item (the block offsets of its lobe counts in sums, then a workgroup scan).
__global__ __launch_bounds__(kFmlsBlock) void fmls_lobe_items(FmlsLaunch L, size_t nv) {
  __shared__ unsigned long long s[kFmlsBlock];
  const size_t slot = (size_t)blockIdx.x * kFmlsBlock + threadIdx.x;
  const unsigned long long c = slot < nv ? (unsigned long long)L.work.nlobes[slot] : 0ull;
  s[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < kFmlsBlock; o <<= 1) {
    const unsigned long long x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  if (slot >= nv) return;
  const unsigned long long off = L.sums[blockIdx.x] + s[threadIdx.x] - c;
  for (unsigned long long l = 0; l < c; ++l) L.work.items[off + l] = (uint32_t)((slot << 11) | l);
}

// Lobe revision over the chunk's lobe items, one lane per lobe (so a voxel of many lobes spreads
// over many lanes instead of holding back its wavefront), the lanes of the grid striding over the
// items.
template <int LMAX>
__global__ __launch_bounds__(kFmlsBlock) void fmls_revise_kernel(FmlsDeviceSet S, FmlsLaunch L, size_t v0) {
  // per-lane Legendre values of the Newton search: lane t's column at s_al[t + kFmlsBlock k]
  __shared__ double s_al[(LMAX / 2 + 1) * (LMAX / 2 + 1) * kFmlsBlock];
  const unsigned long long n = *L.work.nitems;
  const size_t stride = (size_t)gridDim.x * kFmlsBlock;
  for (size_t t = (size_t)blockIdx.x * kFmlsBlock + threadIdx.x; t < n; t += stride) {
    const uint32_t it = L.work.items[t];
    const size_t slot = it >> 11;
    const Tabs T = tabs_of(L.work, slot, L.chunk);
    fmls_revise_lobe<LMAX>(S, T, (int)(it & 0x7ffu), L.fod + (v0 + slot) * (size_t)L.ld_fod, s_al + threadIdx.x);
  }
}

// Each voxel's kept lobes in order, at most maxnum of them (fod2fixel.cpp:139-150): its fixel
// count and summary.
__global__ __launch_bounds__(kFmlsBlock) void fmls_finalize_kernel(FmlsDeviceSet S, FmlsLaunch L, size_t v0,
                                                                    size_t nv) {
  const size_t slot = (size_t)blockIdx.x * kFmlsBlock + threadIdx.x;
  if (slot >= nv) return;
  const size_t vox = v0 + slot;
  const Tabs T = tabs_of(L.work, slot, L.chunk);
  const int nl = L.work.nlobes[slot];
  int nk = 0;
  double sp0 = 0.0, sp1 = 0.0, sx = NAN, sy = NAN, sz = NAN;
  for (int l = 0; l < nl; ++l) {
    if (T.np[l] < 0) continue;
    if (S.maxnum > 0 && nk == S.maxnum) break;
    if (nk == 0) {
      sp0 = T.peak[l];
      sx = T.mean[3 * l];
      sy = T.mean[3 * l + 1];
      sz = T.mean[3 * l + 2];
    } else if (nk == 1) {
      sp1 = T.peak[l];
    }
    ++nk;
  }
  L.count[vox] = (uint32_t)nk;
  if (L.summary) {
    float* o = L.summary + vox * 5;
    o[0] = (float)sp0;
    o[1] = (float)sp1;
    o[2] = (float)sx;
    o[3] = (float)sy;
    o[4] = (float)sz;
  }
}

// fixels of each block of kFmlsBlock voxels of the chunk
__global__ __launch_bounds__(kFmlsBlock) void fmls_block_sums(const uint32_t* __restrict__ count, size_t v0, size_t nv,
                                                               unsigned long long* __restrict__ sums) {
  __shared__ unsigned long long s[kFmlsBlock];
  const size_t v = (size_t)blockIdx.x * kFmlsBlock + threadIdx.x;
  s[threadIdx.x] = v < nv ? (unsigned long long)count[v0 + v] : 0ull;
  __syncthreads();
  for (int o = kFmlsBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// exclusive scan of the block sums in place, starting at the fixels of the earlier chunks
// (*total), which then becomes the running total including this chunk (one workgroup)
__global__ __launch_bounds__(kFmlsBlock) void fmls_scan_sums(unsigned long long* sums, size_t nb,
                                                              unsigned long long* total) {
  __shared__ unsigned long long s[kFmlsBlock];
  unsigned long long carry = *total;
  for (size_t b = 0; b < nb; b += kFmlsBlock) {
    const size_t i = b + threadIdx.x;
    const unsigned long long v = i < nb ? sums[i] : 0ull;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kFmlsBlock; o <<= 1) {
      const unsigned long long x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
      __syncthreads();
      s[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
    carry += s[kFmlsBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// each voxel's offset (0 for voxels without fixels, as fod2fixel's index image) and its kept
// lobes written there, as far as the capacity goes
__global__ __launch_bounds__(kFmlsBlock) void fmls_emit(FmlsLaunch L, size_t v0, size_t nv) {
  __shared__ unsigned long long s[kFmlsBlock];
  const size_t slot = (size_t)blockIdx.x * kFmlsBlock + threadIdx.x;
  const unsigned long long c = slot < nv ? (unsigned long long)L.count[v0 + slot] : 0ull;
  s[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < kFmlsBlock; o <<= 1) {
    const unsigned long long x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  if (slot >= nv) return;
  const unsigned long long off = L.sums[blockIdx.x] + s[threadIdx.x] - c;
  L.offset[v0 + slot] = c ? (uint32_t)off : 0u;
  if (!c) return;
  const Tabs T = tabs_of(L.work, slot, L.chunk);
  const int nl = L.work.nlobes[slot];
  unsigned long long k = 0;
  for (int l = 0; l < nl && k < c; ++l) {
    if (T.np[l] < 0) continue;
    const unsigned long long q = off + k;
    if (q < L.capacity) {
      float* o = L.fixels + q * 5;
      o[0] = (float)T.mean[3 * l];
      o[1] = (float)T.mean[3 * l + 1];
      o[2] = (float)T.mean[3 * l + 2];
      o[3] = (float)T.integ[l];
      o[4] = (float)T.peak[l];
    }
    ++k;
  }
}

// direction a (amplitude ka) is visited before direction b (amplitude kb): higher amplitude, or
// the same amplitude and a lower index -- a strict order on the positive directions, so any
// correct sort yields the reference's visiting order
__device__ __forceinline__ bool visited_before(double ka, int a, double kb, int b) {
  return (ka > kb) | ((ka == kb) & (a < b));   // no short circuit: compares and selects, no branches
}

// Bitonic sort into visiting order of 64 x PER (amplitude, direction) entries, entry e = PER lane
// + k in key[k], idx[k] of lane `lane`: compare-exchanges at distances below PER within the lane's
// registers, at PER and above between lanes (partner lane ^ (j / PER)).
template <int PER>
__device__ __forceinline__ void fmls_bitonic(double (&key)[PER], int (&idx)[PER], int lane) {
  for (int sz = 2; sz <= 64 * PER; sz <<= 1) {
    for (int j = sz >> 1; j >= PER; j >>= 1) {
      const int lj = j / PER;
      const bool first = ((lane & lj) == 0) == (((lane * PER) & sz) == 0);   // keep the earlier entry
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const double ok2 = __shfl_xor(key[k], lj);
        const int oi = __shfl_xor(idx[k], lj);
        const bool take = visited_before(ok2, oi, key[k], idx[k]) == first;
        key[k] = take ? ok2 : key[k];
        idx[k] = take ? oi : idx[k];
      }
    }
#pragma unroll
    for (int j = PER / 2; j > 0; j >>= 1) {
      if (j >= sz) continue;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        if (k & j) continue;
        const int k2 = k | j;
        const bool up = (((lane * PER) | k) & sz) == 0;
        // ascending: k2's entry moves to k when it is visited first; descending: when it is not
        // (equal entries are padding, where a swap changes nothing)
        const bool sw = visited_before(key[k2], idx[k2], key[k], idx[k]) == up;
        const double a0 = key[k], a1 = key[k2];
        const int i0 = idx[k], i1 = idx[k2];
        key[k] = sw ? a1 : a0;
        key[k2] = sw ? a0 : a1;
        idx[k] = sw ? i1 : i0;
        idx[k2] = sw ? i0 : i1;
      }
    }
  }
}

// The sorted entries [0, cnt): their amplitudes to rev and their positions to s_pos (LDS, ndir
// entries), then each direction with its mask of neighbours visited before it to ordm.
template <int PER>
__device__ __forceinline__ void fmls_sorted_out(const FmlsDeviceSet& S, const FmlsLaunch& L, size_t slot, int lane,
                                                int cnt, const double (&key)[PER], const int (&idx)[PER],
                                                uint16_t* s_pos) {
  for (int d = lane; d < S.ndir; d += 64) s_pos[d] = 0xffff;
  __syncthreads();
  const size_t cs = L.chunk;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = lane * PER + k;
    if (e < cnt) {
      s_pos[idx[k]] = (uint16_t)e;
      L.work.rev[(size_t)e * cs + slot] = key[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = lane * PER + k;
    if (e < cnt) {
      const AdjRow ar = adj_row(S, idx[k]);   // one 16-byte load per direction
      unsigned m = 0;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int j = (int16_t)((ar.w[r / 2] >> (16 * (r % 2))) & 0xffffu);
        m |= (j >= 0 && (int)s_pos[j < 0 ? 0 : j] < e) ? 1u << r : 0u;
      }
      L.work.ordm[(size_t)e * cs + slot] = (uint32_t)idx[k] | (m << 16);
    }
  }
}

constexpr int kSortPer = 32;        // entries per lane of the full sort: 64 x 32 = 2048 >= kFmlsMaxDirs
constexpr int kSortSmallPer = 16;   // of the sort of at most 1024 positive directions (nearly every voxel)

// Workgroup b of a grid of n, placed so that each of the 8 XCDs (which take workgroups round-robin)
// runs one contiguous range of them: neighbouring voxels then write the same lines of the
// lane-interleaved arrays through one L2.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned b, unsigned n) {
  const unsigned x = b % 8u, q = n / 8u, r = n % 8u;
  return x * q + (x < r ? x : r) + b / 8u;
}

// One wavefront per voxel of the chunk: the check that the largest |amplitude| (the lowest index
// among equal ones) is positive (fmls.cpp:171-184), then the positive directions in visiting order
// by a bitonic sort of (amplitude, direction) entries held in registers, padded with entries
// visited after every positive one: with at most 1024 positive directions (nearly every voxel)
// they are first packed through LDS and sorted as 64 x 16 entries, otherwise as 64 x 32.  Writes
// npos (0: nothing to segment), ordm (the directions and which of their neighbours are visited
// before them, from the directions' positions in an LDS table) and rev (the amplitudes).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void fmls_sort_kernel(
    FmlsDeviceSet S, FmlsLaunch L, size_t v0, size_t nv) {
  // the small sort's staging area (keys, then directions), then the visiting position of each
  // direction (0xffff: not visited)
  __shared__ double s_buf[64 * kSortSmallPer * 10 / 8];
  const int ndir = S.ndir;
  const size_t slot = xcd_contiguous(blockIdx.x, gridDim.x);
  if (slot >= nv) return;
  const int lane = threadIdx.x;
  const size_t vox = v0 + slot;
  const double c0 = (double)L.fod[vox * (size_t)L.ld_fod];
  const bool ok = (L.valid == nullptr || L.valid[vox] != 0) && c0 > 0.0 && isfinite(c0);
  if (!ok) {
    if (lane == 0) L.work.npos[slot] = 0;
    return;
  }
  const double* a = L.work.amps + slot * (size_t)fmls_amp_ld(ndir);
  double key[kSortPer];
  int idx[kSortPer];
  double best = -1.0;   // largest |amplitude| (NaN never wins) and its lowest index
  int bi = 0;
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < kSortPer; ++k) {
    const int d = k * 64 + lane;   // ascending in k: the first of equal maxima is kept
    const double v = d < ndir ? a[d] : 0.0;
    if (d < ndir && fabs(v) > best) {
      best = fabs(v);
      bi = d;
    }
    const bool pos = d < ndir && v > 0.0;
    key[k] = pos ? v : 0.0;
    idx[k] = pos ? d : 0xffff;
    cnt += pos ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o);
    const int oi = __shfl_xor(bi, o);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
    cnt += __shfl_xor(cnt, o);
  }
  if (!(a[bi] > 0.0)) {   // bi is the same on every lane
    if (lane == 0) L.work.npos[slot] = 0;
    return;
  }
  uint16_t* s_pos = reinterpret_cast<uint16_t*>(s_buf);
  if (cnt <= 64 * kSortSmallPer) {
    // the positive entries packed into staging slots [0, cnt) (in direction order), then sorted
    // as 64 x 16 entries
    double* s_key = s_buf;
    uint16_t* s_idx = reinterpret_cast<uint16_t*>(s_buf + 64 * kSortSmallPer);
    int base = 0;
#pragma unroll
    for (int k = 0; k < kSortPer; ++k) {
      const bool pos = idx[k] != 0xffff;
      const unsigned long long m = __ballot(pos);
      if (pos) {
        const int o = base + (int)__popcll(m & ((1ull << lane) - 1ull));
        s_key[o] = key[k];
        s_idx[o] = (uint16_t)idx[k];
      }
      base += (int)__popcll(m);
    }
    __syncthreads();
    double ks[kSortSmallPer];
    int is[kSortSmallPer];
#pragma unroll
    for (int k = 0; k < kSortSmallPer; ++k) {
      const int e = lane * kSortSmallPer + k;
      ks[k] = e < cnt ? s_key[e] : 0.0;
      is[k] = e < cnt ? (int)s_idx[e] : 0xffff;
    }
    __syncthreads();   // the staging area becomes s_pos
    fmls_bitonic<kSortSmallPer>(ks, is, lane);
    fmls_sorted_out<kSortSmallPer>(S, L, slot, lane, cnt, ks, is, s_pos);
  } else {
    fmls_bitonic<kSortPer>(key, idx, lane);
    fmls_sorted_out<kSortPer>(S, L, slot, lane, cnt, key, idx, s_pos);
  }
  if (lane == 0) L.work.npos[slot] = cnt;
}

// Lobe growth for the voxels of the chunk, one lane per voxel, `lanes` voxels per workgroup, each
// lane's labels in its own LDS column (so the neighbour-label reads of the growth never leave the
// CU).  The final labels go to the work arrays for the per-lobe kernel; nlobes[slot] = the lobe
// count, or -1 when the labels could not number the voxel's lobes.  redo: only the voxels left at
// -1 by an earlier pass are grown.
template <class Lab>
__global__ __launch_bounds__(64) void fmls_grow_kernel(FmlsDeviceSet S, FmlsLaunch L, size_t nv, int lanes, int redo) {
  const int lane = threadIdx.x;
  const size_t slot = (size_t)blockIdx.x * lanes + lane;
  if (lane >= lanes || slot >= nv) return;
  if (redo && L.work.nlobes[slot] != -1) return;
  const int P = L.work.npos[slot];
  int nl = 0;
  if (P > 0) {
    const Tabs T = tabs_of(L.work, slot, L.chunk);
    const Lab lab{(uint32_t)lane, (uint32_t)lanes};
    nl = fmls_grow_voxel(S, T, lab, P);
    if (nl >= 0)
      for (int d = 0; d < S.ndir; ++d) T.lab[d] = (int16_t)lab.get(d);
  }
  L.work.nlobes[slot] = nl;
}

// ---- HIP launches ----

constexpr int kTileD = 128, kTileV = 64, kTileK = 16;

// amps[v * lda + d] for the nv voxels of the chunk starting at v0
__global__ __launch_bounds__(256) void fmls_amp_kernel(const double* __restrict__ sht, int ndir, int nsh,
                                                       const float* __restrict__ fod, int ld, size_t v0, size_t nv,
                                                       size_t lda, double* __restrict__ amps) {
  __shared__ double sA[kTileK][kTileD + 1];
  __shared__ double sB[kTileK][kTileV + 1];
  const int tid = threadIdx.x;
  const int d0 = blockIdx.x * kTileD;
  const size_t vb = (size_t)blockIdx.y * kTileV;
  const int ty = tid / 16, tx = tid % 16;   // directions ty*8 .. ty*8+7, voxels tx + 16 i
  double acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < nsh; k0 += kTileK) {
    for (int i = tid; i < kTileD * kTileK; i += 256) {
      const int dl = i / kTileK, kk = i % kTileK;
      const int d = d0 + dl, k = k0 + kk;
      sA[kk][dl] = (d < ndir && k < nsh) ? sht[(size_t)d * nsh + k] : 0.0;
    }
    for (int i = tid; i < kTileV * kTileK; i += 256) {
      const int vl = i / kTileK, kk = i % kTileK;
      const size_t v = vb + vl;
      const int k = k0 + kk;
      sB[kk][vl] = (v < nv && k < nsh) ? (double)fod[(v0 + v) * (size_t)ld + k] : 0.0;
    }
    __syncthreads();
    const int kend = nsh - k0 < kTileK ? nsh - k0 : kTileK;
    for (int kk = 0; kk < kend; ++kk) {
      double a[8], b[4];
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = sA[kk][ty * 8 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i) b[i] = sB[kk][tx + 16 * i];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = fma(a[j], b[i], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const size_t v = vb + tx + 16 * i;
    if (v >= nv) continue;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int d = d0 + ty * 8 + j;
      if (d < ndir) amps[v * lda + d] = acc[i][j];
    }
  }
}

__global__ void fmls_reset_total(unsigned long long* total) {
  if (threadIdx.x == 0) *total = 0ull;
}

// the lobe revision's grid: workgroups striding over the items, a few per SIMD of the largest
// MI355X part, fewer when the pass is small
constexpr unsigned kReviseGrid = 8192;

template <int LMAX>
hipError_t launch_revise(const FmlsDeviceSet& S, const FmlsLaunch& L, size_t v0, size_t nv, hipStream_t stream) {
  const size_t nb = (nv * 8 + kFmlsBlock - 1) / kFmlsBlock;   // about 8 lobes per voxel and lane
  const unsigned grid = (unsigned)(nb < kReviseGrid ? (nb ? nb : 1) : kReviseGrid);
  fmls_revise_kernel<LMAX><<<dim3(grid), dim3(kFmlsBlock), 0, stream>>>(S, L, v0);
  return hipGetLastError();
}

}  // namespace

hipError_t fmls_launch(const FmlsDeviceSet& S, const FmlsLaunch& L, hipStream_t stream) {
  if (S.ndir < 64 || S.ndir > kFmlsMaxDirs || L.chunk == 0 || L.chunk % kFmlsBlock || L.chunk > (1u << 17))
    return hipErrorInvalidValue;
  if (S.nsh != (S.lmax + 1) * (S.lmax + 2) / 2 || L.ld_fod < S.nsh) return hipErrorInvalidValue;
  if (S.lanes8 < 1 || S.lanes8 > 64 || S.lanes16 < 1 || S.lanes16 > 64) return hipErrorInvalidValue;
  hipError_t e;
  fmls_reset_total<<<1, 64, 0, stream>>>(L.total);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (size_t v0 = 0; v0 < L.nvox; v0 += L.chunk) {
    const size_t nv = L.nvox - v0 < L.chunk ? L.nvox - v0 : L.chunk;
    const unsigned nb = (unsigned)((nv + kFmlsBlock - 1) / kFmlsBlock);
    const dim3 grid((unsigned)((S.ndir + kTileD - 1) / kTileD), (unsigned)((nv + kTileV - 1) / kTileV));
    fmls_amp_kernel<<<grid, dim3(256), 0, stream>>>(S.sht, S.ndir, S.nsh, L.fod, L.ld_fod, v0, nv,
                                                    (size_t)fmls_amp_ld(S.ndir), L.work.amps);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_sort_kernel<<<dim3((unsigned)nv), dim3(64), 0, stream>>>(S, L, v0, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // growth with 8-bit labels, then the voxels of more than 127 lobes again with 16-bit ones
    fmls_grow_kernel<Lab8><<<dim3((unsigned)((nv + S.lanes8 - 1) / S.lanes8)), dim3(64),
                             (size_t)S.ndir * S.lanes8, stream>>>(S, L, nv, S.lanes8, 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_grow_kernel<Lab16><<<dim3((unsigned)((nv + S.lanes16 - 1) / S.lanes16)), dim3(64),
                              (size_t)S.ndir * S.lanes16 * 2, stream>>>(S, L, nv, S.lanes16, 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the lobes as items: block sums of the lobe counts, their scan, the items
    fmls_block_sums<<<dim3(nb), dim3(kFmlsBlock), 0, stream>>>(reinterpret_cast<const uint32_t*>(L.work.nlobes), 0,
                                                               nv, L.sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_reset_total<<<1, 64, 0, stream>>>(L.work.nitems);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_scan_sums<<<1, dim3(kFmlsBlock), 0, stream>>>(L.sums, nb, L.work.nitems);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_lobe_items<<<dim3(nb), dim3(kFmlsBlock), 0, stream>>>(L, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    switch (S.lmax) {
      case 2: e = launch_revise<2>(S, L, v0, nv, stream); break;
      case 4: e = launch_revise<4>(S, L, v0, nv, stream); break;
      case 6: e = launch_revise<6>(S, L, v0, nv, stream); break;
      case 8: e = launch_revise<8>(S, L, v0, nv, stream); break;
      case 10: e = launch_revise<10>(S, L, v0, nv, stream); break;
      case 12: e = launch_revise<12>(S, L, v0, nv, stream); break;
      default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    fmls_finalize_kernel<<<dim3(nb), dim3(kFmlsBlock), 0, stream>>>(S, L, v0, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_block_sums<<<dim3(nb), dim3(kFmlsBlock), 0, stream>>>(L.count, v0, nv, L.sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_scan_sums<<<1, dim3(kFmlsBlock), 0, stream>>>(L.sums, nb, L.total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fmls_emit<<<dim3(nb), dim3(kFmlsBlock), 0, stream>>>(L, v0, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mrcsd
