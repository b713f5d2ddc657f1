// FOD segmentation on gfx950: fod2fixel's FMLS (src/dwi/fmls.cpp:166-377) for a batch of voxels.
//
// Per chunk of voxels:
//   * fmls_amp_kernel: the FOD amplitudes on the segmentation directions (1281 for fod2fixel) as
//     one FP64 GEMM, amps[v][d] = sum_k SHT[d][k] F_v[k] accumulated by FMA in ascending k,
//     LDS-tiled (128 directions x 64 voxels per workgroup, 8 x 4 outputs per thread);
//   * fmls_segment_kernel<LMAX, false>: one wavefront per voxel (persistent), the segmentation:
//       - the positive directions, sorted by decreasing amplitude (lowest index first among equal
//         amplitudes, the reference's multimap order) by a bitonic sort in LDS;
//       - lobe growth in that order, exactly as the reference: a direction adjacent to no positive
//         lobe seeds one, adjacent to one joins it, adjacent to several either merges them (bridge
//         amplitude over the peak of the highest-numbered adjacent lobe above the merge ratio) or
//         waits, outside every lobe, for its assignment to the lowest-numbered of them after the
//         loop.  Every lane follows this loop on the same LDS data, so the decisions are uniform;
//         lanes 0-7 read the eight neighbours' labels, lane 0 updates the lobe records, and the
//         whole wavefront relabels directions after a merge;
//       - per lobe (a lane each): the integral threshold, the Newton peak search from each of its
//         peaks (Math::SH::get_peak in double with the precomputed Legendre table, SH.h:398-559),
//         a revision kept only when it stays nearest to that peak and inside the lobe (nearest
//         direction through the FastLookupSet grid), the peak threshold, the fixel direction;
//       - the voxel's fixels go to a staging buffer at an atomically reserved offset.
//     Lobe records live in LDS for up to 64 lobes and 512 peaks; a voxel beyond that is queued;
//   * fmls_segment_kernel<LMAX, true>: the queued voxels again with the lobe records in a global
//     workspace sized for the worst case (as many lobes and peaks as directions).
// After the last chunk an exclusive scan of the fixel counts gives each voxel's offset in voxel
// order (the order fod2fixel's index image uses, cmd/fod2fixel.cpp:239-280) and the staged fixels
// are moved there.
// Negative directions only ever form negative lobes, which the reference discards and which take
// no part in positive lobes' adjacency, order, merges or indices, so they are not visited.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fmls.h"
#include "wave_utils.h"

namespace mrcsd {
namespace {

constexpr int kSortMax = 2048;

__device__ __forceinline__ int fm_mpos(int l, int m) { return l * l / 4 + m; }
__device__ __forceinline__ int fm_sh(int l, int m) { return l * (l + 1) / 2 + m; }

// Math::SH::derivatives with PrecomputedAL<double> (SH.h:436-470, 568-647): out = {amplitude,
// d/del, d/daz, d2/del2, d2/deldaz, d2/daz2}.  The interpolated Legendre values go to this lane's
// column of an LDS table (entry k at AL[64 k]); the loops over m are not unrolled, which keeps the
// kernel's register count down for the segmentation loop around it.
template <int LMAX>
__device__ void fmls_derivatives(const double* sh, double el, double az, const double* __restrict__ al, int al_ndir,
                                 double* AL, double* out) {
  constexpr int NAL = (LMAX / 2 + 1) * (LMAX / 2 + 1);
  const double sel = sin(el);
  const bool atpole = sel < 1e-4;
  {
    const double inc = M_PI / (al_ndir - 1);
    double f2 = el / inc, f1;
    int i;
    if (!(f2 >= 0.0)) {             // (also a NaN elevation: the first row, never out of bounds)
      i = 0; f1 = 1.0; f2 = 0.0;
    } else if (f2 >= (double)(al_ndir - 1)) {
      i = al_ndir - 1; f1 = 1.0; f2 = 0.0;
    } else {
      i = (int)f2;
      f2 -= i;
      f1 = 1.0 - f2;
    }
    const double* p1 = al + (size_t)i * NAL;
    const double* p2 = p1 + (f2 != 0.0 ? NAL : 0);
    for (int k = 0; k < NAL; ++k) {
      double v = f1 * p1[k];
      if (f2 != 0.0) v += f2 * p2[k];
      AL[64 * k] = v;
    }
  }
#define FM_AL(l, m) AL[64 * fm_mpos((l), (m))]
  double amplitude = sh[0] * FM_AL(0, 0);
  double dSH_del = 0.0, dSH_daz = 0.0, d2SH_del2 = 0.0, d2SH_deldaz = 0.0, d2SH_daz2 = 0.0;
  for (int l = 2; l <= LMAX; l += 2) {
    const double v = sh[fm_sh(l, 0)];
    amplitude += v * FM_AL(l, 0);
    dSH_del += v * sqrt((double)(l * (l + 1))) * FM_AL(l, 1);
    d2SH_del2 += v * (sqrt((double)(l * (l + 1) * (l - 1) * (l + 2))) * FM_AL(l, 2) - l * (l + 1) * FM_AL(l, 0)) / 2.0;
  }
#pragma unroll 1
  for (int m = 1; m <= LMAX; ++m) {
    const double caz = M_SQRT2 * cos(m * az);
    const double saz = M_SQRT2 * sin(m * az);
    for (int l = (m & 1) ? m + 1 : m; l <= LMAX; l += 2) {
      const double vp = sh[fm_sh(l, m)];
      const double vm = sh[fm_sh(l, -m)];
      amplitude += (vp * caz + vm * saz) * FM_AL(l, m);
      double tmp = sqrt((double)((l + m) * (l - m + 1))) * FM_AL(l, m - 1);
      if (l > m) tmp -= sqrt((double)((l - m) * (l + m + 1))) * FM_AL(l, m + 1);
      tmp /= -2.0;
      dSH_del += (vp * caz + vm * saz) * tmp;
      double tmp2 = -((l + m) * (l - m + 1) + (l - m) * (l + m + 1)) * FM_AL(l, m);
      if (m == 1) tmp2 -= sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * FM_AL(l, 1);
      else tmp2 += sqrt((double)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * FM_AL(l, m - 2);
      if (l > m + 1) tmp2 += sqrt((double)((l - m) * (l + m + 1) * (l - m - 1) * (l + m + 2))) * FM_AL(l, m + 2);
      tmp2 /= 4.0;
      d2SH_del2 += (vp * caz + vm * saz) * tmp2;
      if (atpole) {
        dSH_daz += (vm * caz - vp * saz) * tmp;
      } else {
        d2SH_deldaz += m * (vm * caz - vp * saz) * tmp;
        dSH_daz += m * (vm * caz - vp * saz) * FM_AL(l, m);
        d2SH_daz2 -= (vp * caz + vm * saz) * m * m * FM_AL(l, m);
      }
    }
  }
#undef FM_AL
  if (!atpole) {
    dSH_daz /= sel;
    d2SH_deldaz /= sel;
    d2SH_daz2 /= sel * sel;
  }
  out[0] = amplitude; out[1] = dSH_del; out[2] = dSH_daz;
  out[3] = d2SH_del2; out[4] = d2SH_deldaz; out[5] = d2SH_daz2;
}

// Math::SH::get_peak (SH.h:514-559): Newton steps along the gradient, at most 0.2 rad each,
// until a step is below 1e-4 rad; NaN after 50 steps
template <int LMAX>
__device__ double fmls_get_peak(const double* sh, double* dir, const double* __restrict__ al, int al_ndir, double* AL) {
  const double MAX_DIR_CHANGE = 0.2, ANGLE_TOLERANCE = 1e-4;
  for (int it = 0; it < 50; ++it) {
    const double az = atan2(dir[1], dir[0]);
    const double el = atan2(hypot(dir[0], dir[1]), dir[2]);
    double d[6];
    fmls_derivatives<LMAX>(sh, el, az, al, al_ndir, AL, d);
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
    const double ca = cos(az), sa = sin(az), ce = cos(el), se = sin(el);
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

// Lobe records (FOD_lobe, fmls.h:61-137).  A lobe's peaks are the seeds of the lobes merged into
// it, kept as a linked list through next[seed direction] in the reference's order.
struct LobeTab {
  double* peak;      // [cap] maximal peak amplitude
  double* mean;      // [cap][3] amplitude-weighted direction sum
  double* integ;     // [cap]
  double* rev;       // [rev_cap][3] peak directions during the revision
  int16_t* head;     // [cap] first and last seed of the peak list
  int16_t* tail;
  int16_t* np;       // [cap] peaks; -1 after the revision: lobe dropped
  int cap, rev_cap;
};

// FOD_lobe::add (fmls.h:88-97); called by one lane
__device__ __forceinline__ void lobe_add(const FmlsDeviceSet& S, const LobeTab& T, int l, int d, double v) {
  const double* u = S.unit + 3 * d;
  const double w = S.weights[d];
  double* m = T.mean + 3 * l;
  const double mult = (m[0] * u[0] + m[1] * u[1] + m[2] * u[2]) > 0.0 ? 1.0 : -1.0;
  const double av = fabs(v);
  m[0] += u[0] * mult * av * w;
  m[1] += u[1] * mult * av * w;
  m[2] += u[2] * mult * av * w;
  T.integ[l] += fabs(v * w);
}

// FOD_lobe::merge (fmls.h:112-127): B into A; called by one lane
__device__ __forceinline__ void lobe_merge(const LobeTab& T, int16_t* next, int A, int B) {
  if (T.peak[B] > T.peak[A]) {
    T.peak[A] = T.peak[B];
    next[T.tail[B]] = T.head[A];     // B's peaks first
    T.head[A] = T.head[B];
  } else {
    next[T.tail[A]] = T.head[B];     // B's peaks after A's
    T.tail[A] = T.tail[B];
  }
  T.np[A] = (int16_t)(T.np[A] + T.np[B]);
  double* a = T.mean + 3 * A;
  const double* b = T.mean + 3 * B;
  const double mult = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) > 0.0 ? 1.0 : -1.0;
  const double ib = T.integ[B];
  a[0] += b[0] * ib * mult;
  a[1] += b[1] * ib * mult;
  a[2] += b[2] * ib * mult;
  T.integ[A] += ib;
}

__device__ __forceinline__ void lobe_copy(const LobeTab& T, int dst, int src) {
  T.peak[dst] = T.peak[src];
  T.mean[3 * dst] = T.mean[3 * src];
  T.mean[3 * dst + 1] = T.mean[3 * src + 1];
  T.mean[3 * dst + 2] = T.mean[3 * src + 2];
  T.integ[dst] = T.integ[src];
  T.head[dst] = T.head[src];
  T.tail[dst] = T.tail[src];
  T.np[dst] = T.np[src];
}

__device__ __forceinline__ int wave_sum_int(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__device__ __forceinline__ int wave_excl_scan(int x, int lane) {
  int s = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl(s, lane >= off ? lane - off : lane);
    if (lane >= off) s += t;
  }
  return s - x;
}

template <int LMAX, bool BIG>
__global__ __launch_bounds__(64) void fmls_segment_kernel(FmlsDeviceSet S, FmlsLaunch L, size_t v0, size_t nv) {
  constexpr int NSH = (LMAX + 1) * (LMAX + 2) / 2;
  constexpr int NL = BIG ? 1 : kFmlsMaxLobes;
  __shared__ double s_val[kFmlsMaxDirs];     // amplitudes; peak directions during the revision
  __shared__ uint16_t s_ord[kSortMax];       // positive directions in visiting order
  __shared__ int16_t s_lab[kFmlsMaxDirs];    // lobe of each direction; -1 none; -2 - l: waits for lobe l
  __shared__ int16_t s_next[kFmlsMaxDirs];   // peak lists
  __shared__ double s_sh[NSH];
  __shared__ double s_al[(LMAX / 2 + 1) * (LMAX / 2 + 1) * 64];   // per-lane Legendre values
  __shared__ double s_lpeak[NL], s_lmean[3 * NL], s_lint[NL];
  __shared__ int16_t s_lhead[NL], s_ltail[NL], s_lnp[NL];

  const int lane = threadIdx.x;
  const int ndir = S.ndir;
  LobeTab T;
  if constexpr (BIG) {
    double* base = L.big_ws + (size_t)blockIdx.x * fmls_big_stride(ndir);
    T.peak = base;
    T.mean = base + ndir;
    T.integ = base + 4 * (size_t)ndir;
    T.rev = base + 5 * (size_t)ndir;
    int16_t* h = (int16_t*)(base + 8 * (size_t)ndir);
    T.head = h;
    T.tail = h + ndir;
    T.np = h + 2 * ndir;
    T.cap = ndir;
    T.rev_cap = ndir;
  } else {
    T.peak = s_lpeak;
    T.mean = s_lmean;
    T.integ = s_lint;
    T.rev = s_val;
    T.head = s_lhead;
    T.tail = s_ltail;
    T.np = s_lnp;
    T.cap = kFmlsMaxLobes;
    T.rev_cap = kFmlsMaxDirs / 3;
  }

  for (;;) {
    unsigned long long k = 0;
    if (lane == 0) k = atomicAdd(L.counters + (BIG ? 2 : 0), 1ull);
    k = uniform64(k);
    size_t idx;
    if constexpr (BIG) {
      if (k >= uniform64(L.counters[1])) break;
      idx = (size_t)uniform32((int)L.queue[k]);
    } else {
      if (k >= nv) break;
      idx = k;
    }
    const size_t vox = v0 + idx;
    const float* f = L.fod + vox * (size_t)L.ld_fod;
    const double c0 = (double)f[0];
    bool ok = (L.valid == nullptr || L.valid[vox] != 0) && c0 > 0.0 && isfinite(c0);
    ok = uniform32(ok ? 1 : 0) != 0;
    int P = 0, nl = 0, nseeds = 0;
    bool overflow = false;
    if (ok) {
      // amplitudes; the largest |amplitude| (lowest index among equal ones) must be positive
      // (fmls.cpp:171-184)
      for (int kk = lane; kk < NSH; kk += 64) s_sh[kk] = (double)f[kk];
      const double* a = L.amps + idx * (size_t)ndir;
      double best = -1.0;
      int bi = 0x7fffffff;
      for (int d = lane; d < ndir; d += 64) {
        const double v = a[d];
        s_val[d] = v;
        s_lab[d] = -1;
        if (fabs(v) > best) {
          best = fabs(v);
          bi = d;
        }
      }
      double nb = -best;
      wave_argmin(nb, bi);
      bi = uniform32(bi);
      wave_sync();
      ok = uniform32(s_val[bi] > 0.0 ? 1 : 0) != 0;
    }
    if (ok) {
      // the positive directions in index order, then sorted: amplitude descending, index ascending
      for (int d0 = 0; d0 < ndir; d0 += 64) {
        const int d = d0 + lane;
        const bool pos = d < ndir && s_val[d] > 0.0;
        const unsigned long long bal = __ballot(pos);
        if (pos) s_ord[P + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)d;
        P += __popcll(bal);
      }
      int SN = 64;
      while (SN < P) SN <<= 1;
      for (int t = P + lane; t < SN; t += 64) s_ord[t] = 0xFFFF;
      wave_sync();
      for (int kb = 2; kb <= SN; kb <<= 1) {
        for (int j = kb >> 1; j > 0; j >>= 1) {
          for (int p = lane; p < (SN >> 1); p += 64) {
            const int t = ((p & ~(j - 1)) << 1) | (p & (j - 1));
            const int u = t | j;
            const int x = s_ord[t], y = s_ord[u];
            const double vx = x == 0xFFFF ? -INFINITY : s_val[x];
            const double vy = y == 0xFFFF ? -INFINITY : s_val[y];
            // x visited after y (the padding goes last)
            const bool x_after = vx < vy || (vx == vy && x > y);
            if (((t & kb) == 0) == x_after) {
              s_ord[t] = (uint16_t)y;
              s_ord[u] = (uint16_t)x;
            }
          }
          wave_sync();
        }
      }

      // ---- lobe growth (fmls.cpp:188-249) ----
      for (int i = 0; i < P; ++i) {
        const int d = uniform32(s_ord[i]);
        const double v = s_val[d];
        int lab = -1;
        if (lane < 8) {
          const int j = S.adj[d * 8 + lane];
          lab = j >= 0 ? (int)s_lab[j] : -1;
          if (lab < -1) lab = -1;     // waiting directions belong to no lobe yet
        }
        if (__ballot(lab >= 0) == 0ull) {
          // a new lobe seeded by d
          if (nl == T.cap || nseeds == T.rev_cap) {
            overflow = true;
            break;
          }
          if (lane == 0) {
            const double* u = S.unit + 3 * d;
            const double w = S.weights[d];
            T.peak[nl] = fabs(v);
            T.mean[3 * nl] = u[0] * fabs(v) * w;
            T.mean[3 * nl + 1] = u[1] * fabs(v) * w;
            T.mean[3 * nl + 2] = u[2] * fabs(v) * w;
            T.integ[nl] = fabs(v * w);
            T.head[nl] = (int16_t)d;
            T.tail[nl] = (int16_t)d;
            T.np[nl] = 1;
            s_next[d] = -1;
            s_lab[d] = (int16_t)nl;
          }
          ++nl;
          ++nseeds;
          wave_sync();
          continue;
        }
        int mn = lab >= 0 ? lab : 0x7fffffff, mx = lab;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
          mn = min(mn, __shfl_xor(mn, o));
          mx = max(mx, __shfl_xor(mx, o));
        }
        mn = uniform32(mn);
        mx = uniform32(mx);
        if (mn == mx) {
          if (lane == 0) {
            s_lab[d] = (int16_t)mn;
            lobe_add(S, T, mn, d, v);
          }
          wave_sync();
          continue;
        }
        if (uniform32(fabs(v) / T.peak[mx] > S.merge_ratio ? 1 : 0) == 0) {
          if (lane == 0) s_lab[d] = (int16_t)(-2 - mn);
          wave_sync();
          continue;
        }
        // merge every adjacent lobe into the lowest-numbered one, in ascending order
        int olab[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) olab[r] = __builtin_amdgcn_readlane(lab, r);
        bool dup = false;
#pragma unroll
        for (int r = 0; r < 8; ++r) dup = dup || (r < lane && olab[r] == lab);
        const bool dist = lab >= 0 && !dup;
        const unsigned long long dm = __ballot(dist);
        const int na = __popcll(dm);
        int rank = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r) rank += (dist && ((dm >> r) & 1ull) && olab[r] < lab) ? 1 : 0;
        int orank[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) orank[r] = __builtin_amdgcn_readlane(rank, r);
        const int A = mn;
        if (lane == 0) {
          for (int q = 1; q < na; ++q) {
            int B = A;
#pragma unroll
            for (int r = 0; r < 8; ++r)
              if (((dm >> r) & 1ull) && orank[r] == q) B = olab[r];
            lobe_merge(T, s_next, A, B);
          }
          s_lab[d] = (int16_t)A;
          lobe_add(S, T, A, d, v);
        }
        wave_sync();
        // merged lobes become A, the lobes above them move down over the erased records; the same
        // for the lobes waiting directions are assigned to (fmls.cpp:220-237)
        for (int dd = lane; dd < ndir; dd += 64) {
          const int lb = s_lab[dd];
          const bool wait = lb <= -2;
          int l = wait ? -2 - lb : lb;
          if (l >= 0) {
            bool merged = false;
            int sh = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
              if (((dm >> r) & 1ull) && olab[r] != A) {
                merged = merged || l == olab[r];
                sh += olab[r] < l ? 1 : 0;
              }
            }
            l = merged ? A : l - sh;
            s_lab[dd] = (int16_t)(wait ? -2 - l : l);
          }
        }
        if (lane == 0) {
          int n = nl;
          for (int q = na - 1; q >= 1; --q) {
            int e = A;
#pragma unroll
            for (int r = 0; r < 8; ++r)
              if (((dm >> r) & 1ull) && orank[r] == q) e = olab[r];
            for (int l = e; l < n - 1; ++l) lobe_copy(T, l, l + 1);
            --n;
          }
        }
        nl -= na - 1;
        wave_sync();
      }

      if (!overflow) {
        // retrospective assignments in visiting order (fmls.cpp:251-252)
        for (int i0 = 0; i0 < P; i0 += 64) {
          const int i = i0 + lane;
          const int d = i < P ? (int)s_ord[i] : 0;
          const int lb = i < P ? (int)s_lab[d] : -1;
          unsigned long long wm = __ballot(lb <= -2);
          while (wm) {
            const int r = __ffsll((long long)wm) - 1;
            wm &= wm - 1ull;
            const int dr = __builtin_amdgcn_readlane(d, r);
            const int l = -2 - __builtin_amdgcn_readlane(lb, r);
            if (lane == 0) {
              s_lab[dr] = (int16_t)l;
              lobe_add(S, T, l, dr, s_val[dr]);
            }
          }
        }
        wave_sync();

        // ---- per lobe: integral threshold, peak revision, peak threshold (fmls.cpp:254-292) ----
        int run = 0;
        for (int l0 = 0; l0 < nl; l0 += 64) {
          const int l = l0 + lane;
          const int np = l < nl ? (int)T.np[l] : 0;
          const int poff = run + wave_excl_scan(np, lane);
          run += wave_sum_int(np);
          if (l < nl) {
            double* rv = T.rev + 3 * (size_t)poff;
            int s = T.head[l];
            for (int p = 0; p < np; ++p) {
              rv[3 * p] = S.unit[3 * s];
              rv[3 * p + 1] = S.unit[3 * s + 1];
              rv[3 * p + 2] = S.unit[3 * s + 2];
              s = s_next[s];
            }
            bool keep = T.integ[l] >= S.integral_threshold;
            double maxp = T.peak[l];
            if (keep) {
              for (int p = 0; p < np; ++p) {
                double nd[3] = {rv[3 * p], rv[3 * p + 1], rv[3 * p + 2]};
                const double nvl = fmls_get_peak<LMAX>(s_sh, nd, S.al, S.al_ndir, s_al + lane);
                if (isfinite(nvl) && isfinite(nd[0]) && isfinite(nd[1]) && isfinite(nd[2])) {
                  double max_dp = 0.0;
                  int nearest = np;
                  for (int q = 0; q < np; ++q) {
                    const double dp = fabs(nd[0] * rv[3 * q] + nd[1] * rv[3 * q + 1] + nd[2] * rv[3 * q + 2]);
                    if (dp > max_dp) {
                      max_dp = dp;
                      nearest = q;
                    }
                  }
                  if (nearest == p && s_lab[fmls_select(S, nd)] == l) {
                    rv[3 * p] = nd[0];
                    rv[3 * p + 1] = nd[1];
                    rv[3 * p + 2] = nd[2];
                    if (p == 0) maxp = nvl;
                  }
                }
              }
              keep = !(maxp < S.peak_threshold);
            }
            if (keep) {
              double* m = T.mean + 3 * l;
              if (S.dir_from_peak) {
                m[0] = rv[0];
                m[1] = rv[1];
                m[2] = rv[2];
              } else {
                const double nrm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                m[0] /= nrm;
                m[1] /= nrm;
                m[2] /= nrm;
              }
              T.peak[l] = maxp;
            } else {
              T.np[l] = -1;
            }
          }
        }
        wave_sync();
      }
    }
    if (overflow) {
      // lobes beyond the LDS tables: the workspace pass takes the voxel
      if (lane == 0) {
        const unsigned long long q = atomicAdd(L.counters + 1, 1ull);
        L.queue[q] = (uint32_t)idx;
      }
      continue;
    }
    // the kept lobes in order, at most maxnum of them (fod2fixel.cpp:139-150)
    int nk = 0;
    if (ok)
      for (int l0 = 0; l0 < nl; l0 += 64) nk += __popcll(__ballot(l0 + lane < nl && T.np[l0 + lane] >= 0));
    if (S.maxnum > 0 && nk > S.maxnum) nk = S.maxnum;
    unsigned long long base = 0;
    if (lane == 0 && nk > 0) base = atomicAdd(L.total, (unsigned long long)nk);
    base = uniform64(base);
    double sp0 = 0.0, sp1 = 0.0, sx = NAN, sy = NAN, sz = NAN;
    int pos0 = 0;
    for (int l0 = 0; l0 < nl && ok && pos0 < nk; l0 += 64) {
      const int l = l0 + lane;
      const bool kept = l < nl && T.np[l] >= 0;
      const unsigned long long km = __ballot(kept);
      const int pos = pos0 + __popcll(km & ((1ull << lane) - 1ull));
      if (kept && pos < nk && base + pos < L.capacity) {
        float* o = L.stage + (base + pos) * 5;
        o[0] = (float)T.mean[3 * l];
        o[1] = (float)T.mean[3 * l + 1];
        o[2] = (float)T.mean[3 * l + 2];
        o[3] = (float)T.integ[l];
        o[4] = (float)T.peak[l];
      }
      if (L.summary) {
        // fixels 0 and 1 of the voxel: the first two kept lobes
        const int lr = l < nl ? l : 0;
        const double pk = T.peak[lr];
        const double mx0 = T.mean[3 * lr], my0 = T.mean[3 * lr + 1], mz0 = T.mean[3 * lr + 2];
        if (pos0 == 0 && km) {
          const int r = __ffsll((long long)km) - 1;
          sp0 = lane_bcast(pk, r);
          sx = lane_bcast(mx0, r);
          sy = lane_bcast(my0, r);
          sz = lane_bcast(mz0, r);
        }
        const int want = 1 - pos0;    // place of fixel 1 among this round's kept lobes
        if (want >= 0 && __popcll(km) > want && nk > 1) {
          unsigned long long mm = km;
          for (int w = 0; w < want; ++w) mm &= mm - 1ull;
          sp1 = lane_bcast(pk, __ffsll((long long)mm) - 1);
        }
      }
      pos0 += __popcll(km);
    }
    if (lane == 0) {
      L.count[vox] = (uint32_t)nk;
      L.stage_off[vox] = (uint32_t)base;
      if (L.summary) {
        float* s = L.summary + vox * 5;
        s[0] = (float)sp0;
        s[1] = (float)sp1;
        s[2] = (float)sx;
        s[3] = (float)sy;
        s[4] = (float)sz;
      }
    }
    wave_sync();
  }
}

// ---- HIP launches ----

constexpr int kTileD = 128, kTileV = 64, kTileK = 16;

__global__ __launch_bounds__(256) void fmls_amp_kernel(const double* __restrict__ sht, int ndir, int nsh,
                                                       const float* __restrict__ fod, int ld, size_t v0, size_t nv,
                                                       double* __restrict__ amps) {
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
      if (d < ndir) amps[v * ndir + d] = acc[i][j];
    }
  }
}

__global__ void fmls_reset_kernel(unsigned long long* counters, unsigned long long* total, int reset_total) {
  if (threadIdx.x < 4) counters[threadIdx.x] = 0ull;
  if (threadIdx.x == 0 && reset_total) *total = 0ull;
}

constexpr int kScanThreads = 256, kScanPer = 8;

__global__ __launch_bounds__(256) void fmls_scan_local(const uint32_t* __restrict__ count, size_t n,
                                                        uint32_t* __restrict__ sums) {
  __shared__ uint32_t s[kScanThreads];
  const size_t b0 = (size_t)blockIdx.x * (kScanThreads * kScanPer) + (size_t)threadIdx.x * kScanPer;
  uint32_t t = 0;
  for (int i = 0; i < kScanPer; ++i)
    if (b0 + i < n) t += count[b0 + i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int o = kScanThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// exclusive scan of the block sums in place, their total after them (one workgroup)
__global__ __launch_bounds__(256) void fmls_scan_sums(uint32_t* sums, size_t nb) {
  __shared__ uint32_t s[kScanThreads];
  uint32_t carry = 0;
  for (size_t b = 0; b < nb; b += kScanThreads) {
    const size_t i = b + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
      const uint32_t x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
      __syncthreads();
      s[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
    carry += s[kScanThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

// offsets in voxel order (0 for voxels without fixels, as fod2fixel's index image) and the fixels
// moved there from the staging buffer
__global__ __launch_bounds__(256) void fmls_compact(const uint32_t* __restrict__ count,
                                                     const uint32_t* __restrict__ stage_off,
                                                     const float* __restrict__ stage, size_t capacity,
                                                     const uint32_t* __restrict__ sums, size_t n,
                                                     uint32_t* __restrict__ offset, float* __restrict__ fixels) {
  __shared__ uint32_t s[kScanThreads];
  const size_t b0 = (size_t)blockIdx.x * (kScanThreads * kScanPer) + (size_t)threadIdx.x * kScanPer;
  uint32_t t = 0;
  for (int i = 0; i < kScanPer; ++i)
    if (b0 + i < n) t += count[b0 + i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {
    const uint32_t x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  size_t off = (size_t)sums[blockIdx.x] + s[threadIdx.x] - t;
  for (int i = 0; i < kScanPer; ++i) {
    const size_t v = b0 + i;
    if (v >= n) break;
    const uint32_t c = count[v];
    offset[v] = c ? (uint32_t)off : 0u;
    const size_t src = stage_off[v];
    if (c && off + c <= capacity && src + c <= capacity)
      for (uint32_t q = 0; q < 5 * c; ++q) fixels[off * 5 + q] = stage[src * 5 + q];
    off += c;
  }
}

template <int LMAX>
hipError_t launch_segment(const FmlsDeviceSet& S, const FmlsLaunch& L, size_t v0, size_t nv, hipStream_t stream) {
  int dev = 0, ncu = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  int per_cu = 0;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fmls_segment_kernel<LMAX, false>, 64, 0)) != hipSuccess)
    return e;
  if (per_cu < 1) per_cu = 1;
  size_t grid = (size_t)ncu * (size_t)per_cu;
  if (grid > nv) grid = nv;
  fmls_segment_kernel<LMAX, false><<<dim3((unsigned)grid), dim3(64), 0, stream>>>(S, L, v0, nv);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the queued voxels; the workgroups find an empty queue and return when nothing overflowed
  fmls_segment_kernel<LMAX, true><<<dim3((unsigned)L.big_blocks), dim3(64), 0, stream>>>(S, L, v0, nv);
  return hipGetLastError();
}

}  // namespace

bool fmls_gpu_enabled() { return false; }

hipError_t fmls_launch(const FmlsDeviceSet& S, const FmlsLaunch& L, hipStream_t stream) {
  if (!fmls_gpu_enabled()) return hipErrorNotSupported;
  if (S.ndir < 64 || S.ndir > kFmlsMaxDirs || L.chunk == 0 || L.big_blocks < 1) return hipErrorInvalidValue;
  if (S.nsh != (S.lmax + 1) * (S.lmax + 2) / 2 || L.ld_fod < S.nsh) return hipErrorInvalidValue;
  hipError_t e;
  for (size_t v0 = 0; v0 < L.nvox; v0 += L.chunk) {
    const size_t nv = L.nvox - v0 < L.chunk ? L.nvox - v0 : L.chunk;
    fmls_reset_kernel<<<1, 64, 0, stream>>>(L.counters, L.total, v0 == 0 ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 grid((unsigned)((S.ndir + kTileD - 1) / kTileD), (unsigned)((nv + kTileV - 1) / kTileV));
    fmls_amp_kernel<<<grid, dim3(256), 0, stream>>>(S.sht, S.ndir, S.nsh, L.fod, L.ld_fod, v0, nv, L.amps);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    switch (S.lmax) {
      case 2: e = launch_segment<2>(S, L, v0, nv, stream); break;
      case 4: e = launch_segment<4>(S, L, v0, nv, stream); break;
      case 6: e = launch_segment<6>(S, L, v0, nv, stream); break;
      case 8: e = launch_segment<8>(S, L, v0, nv, stream); break;
      case 10: e = launch_segment<10>(S, L, v0, nv, stream); break;
      case 12: e = launch_segment<12>(S, L, v0, nv, stream); break;
      default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
  }
  if (L.nvox == 0) {
    fmls_reset_kernel<<<1, 64, 0, stream>>>(L.counters, L.total, 1);
    return hipGetLastError();
  }
  const size_t nb = fmls_scan_blocks(L.nvox);
  fmls_scan_local<<<dim3((unsigned)nb), dim3(kScanThreads), 0, stream>>>(L.count, L.nvox, L.scan_ws);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  fmls_scan_sums<<<1, kScanThreads, 0, stream>>>(L.scan_ws, nb);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  fmls_compact<<<dim3((unsigned)nb), dim3(kScanThreads), 0, stream>>>(L.count, L.stage_off, L.stage, L.capacity,
                                                                      L.scan_ws, L.nvox, L.offset, L.fixels);
  return hipGetLastError();
}

}  // namespace mrcsd
