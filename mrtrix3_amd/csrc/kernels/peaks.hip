// sh2peaks on gfx950: the peaks of an SH series in every voxel (cmd/sh2peaks.cpp:160-247), each
// found by Math::SH::get_peak's Gauss-Newton ascent (core/math/SH.h:514-559) from every search
// direction.
//
// Two kernels per chunk of voxels:
//   * peaks_search_kernel: every (voxel, search direction) pair is a work item; each lane takes the
//     next item as soon as its current search ends (converged, or 50 iterations), so lanes stay busy
//     whatever the spread of iteration counts (a wavefront per voxel would wait for its slowest
//     search: about a third of the lanes' time).  A wavefront reserves 256 items per atomic and
//     hands them to its idle lanes.  Results go to a work buffer: amplitude, direction.
//   * peaks_select_kernel: one wavefront per voxel (a fixed stride over the voxels, no counter:
//     one atomic per voxel on one address serialises) reads its searches' results and collects the
//     accepted peaks in seed order as the reference does (a peak within acos(0.99) of an accepted
//     one is dropped, then the threshold applies), one accepted peak per lane, then chooses the
//     reported ones by decreasing amplitude or by closeness to given directions.
//
// Arithmetic follows the reference's: single precision, with the steps it takes in double where
// a double literal or constant enters (Plm_sph_helper, the /2.0, /-2.0, /4.0 of the derivative
// terms, Math::sqrt2, the 2.0 * daz * del term, the constants of get_peak), and no fused
// multiply-adds (contraction is off in this file).  Results differ from it in the last bits
// through cheaper evaluations of the same quantities: the sines and cosines of the direction's
// angles from its components (not atan2 then sin / cos), cos / sin(m az) by a recurrence, the
// Legendre seed helper with folded constants and its square root in single precision.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>

#include "peaks.h"
#include "wave_utils.h"

#pragma clang fp contract(off)

namespace mrcsd {
namespace {

constexpr double kMaxDirChange = 0.2;    // SH.h:23
constexpr double kAngleTolerance = 1e-4; // SH.h:24
constexpr double kDotThreshold = 0.99;   // sh2peaks.cpp:25
constexpr double kSqrt2 = 1.41421356237309504880;   // Math::sqrt2 (core/math/math.h:43)
constexpr double kPi = 3.14159265358979323846;   // Math::pi (core/math/math.h)

constexpr int nforl(int L) { return (L + 1) * (L + 2) / 2; }
constexpr int nforl_mpos(int L) { return (L / 2 + 1) * (L / 2 + 1); }
constexpr int sh_index(int l, int m) { return l * (l + 1) / 2 + m; }   // SH.h:52-55
constexpr int index_mpos(int l, int m) { return l * l / 4 + m; }     // SH.h:64-67

// legendre.h:66-69 with m = M2 / 2, in double as the reference calls it, with each level's
// (M2 - 1) / M2 folded into one constant at compile time instead of a division per level (the
// result differs from the reference's in the last bits of the double only)
template <int M2>
__device__ __forceinline__ double plm_helper(double x) {
  if constexpr (M2 < 1) {
    return 1.0;
  } else {
    constexpr double f = ((double)M2 - 1.0) / (double)M2;
    return x * f * plm_helper<M2 - 2>(x);
  }
}

__device__ __forceinline__ float fbits(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// a / f for a divisor known at compile time: y = 1 / f folds to the correctly rounded reciprocal,
// and one product with two fused corrections (Markstein's theorem) gives the correctly rounded
// quotient in three instructions instead of a division sequence.  Checked against a / f for every
// float a and every divisor of the recursion below up to lmax 12: equal (zeros' signs included)
// for all finite |a| >= 2^-100; below, where the residual leaves the normal range, it can differ
// by one unit of the last place.
__device__ __forceinline__ float div_const(float a, float f) {
  const float y = 1.0f / f;
  const float q0 = a * y;
  const float r = __builtin_fmaf(f, q0, -a);
  return __builtin_fmaf(-r, y, q0);
}

// legendre.h:129-154 for order M, written into AL at index_mpos(l, M) for even l (SH.h:585-590)
template <int LMAX, int M>
__device__ __forceinline__ void plm_order(float (&AL)[nforl_mpos(LMAX)], float x, float sin2) {
  float buf[LMAX + 1];
  buf[M] = (float)0.282094791773878;
  if constexpr (M > 0)   // the product in double as the reference forms it, its root in single precision
    buf[M] = buf[M] * sqrtf((float)((double)(float)(2 * M + 1) * plm_helper<2 * M>(1.0 * (double)sin2)));
  if constexpr ((M & 1) != 0) buf[M] = -buf[M];
  if constexpr (LMAX > M) {
    float f = sqrtf((float)(2 * M + 3));
    buf[M + 1] = x * f * buf[M];
#pragma unroll
    for (int n = M + 2; n <= LMAX; n++) {
      buf[n] = x * buf[n - 1] - div_const(buf[n - 2], f);
      f = sqrtf((float)(4 * n * n - 1) / (float)(n * n - M * M));
      buf[n] *= f;
    }
  }
  const bool pole = M > 0 && sin2 == 0.0f;   // legendre.h:138-142
#pragma unroll
  for (int l = (M & 1) ? M + 1 : M; l <= LMAX; l += 2) AL[index_mpos(l, M)] = pole ? 0.0f : buf[l];
}

template <int LMAX, int... M>
__device__ __forceinline__ void plm_all(float (&AL)[nforl_mpos(LMAX)], float x, float sin2,
                                        std::integer_sequence<int, M...>) {
  (plm_order<LMAX, M>(AL, x, sin2), ...);
}

// SH.h:568-647: amplitude and derivatives at the direction whose elevation and azimuth have the
// given sines and cosines; c = the SH coefficients.  sqrt2 cos(m az) and sqrt2 sin(m az) for
// m >= 2 come from the m = 0 and m = 1 values by the recurrence X(m+1) = 2 cos(az) X(m) - X(m-1)
// in double (its error, below m^2 double ulps, leaves the float values as the directly evaluated
// ones but for rounding ties) instead of sixteen sin / cos calls.
// SH.h:436-470 (PrecomputedAL::set / get, the -fast option): the values at elevation el by linear
// interpolation between the two rows of tab ([kPeaksFastRows][nforl_mpos(LMAX)], row n at elevation
// n pi / 511) around it; the first or last row alone outside [0, pi).
template <int LMAX>
__device__ __forceinline__ void fast_al(float (&AL)[nforl_mpos(LMAX)], const float* __restrict__ tab, float el) {
  const float inc = (float)(kPi / (kPeaksFastRows - 1));
  float f2 = el / inc, f1;
  int i;
  // the reference's int(f2) < 0 and >= 511 tests, decided on the float before it is converted (a
  // NaN elevation, from an SH series with infinite coefficients, takes the first row as x86's
  // conversion of NaN to INT_MIN makes the reference's host code do)
  if (!(f2 > -1.0f)) {
    i = 0; f1 = 1.0f; f2 = 0.0f;
  } else if (f2 >= (float)(kPeaksFastRows - 1)) {
    i = kPeaksFastRows - 1; f1 = 1.0f; f2 = 0.0f;
  } else {
    i = (int)f2;
    f2 -= (float)i;
    f1 = (float)(1.0 - (double)f2);
  }
  const float* p1 = tab + i * nforl_mpos(LMAX);
  if (f2 != 0.0f) {
    const float* p2 = p1 + nforl_mpos(LMAX);
#pragma unroll
    for (int k = 0; k < nforl_mpos(LMAX); ++k) AL[k] = __fmul_rn(f1, p1[k]) + __fmul_rn(f2, p2[k]);
  } else {
#pragma unroll
    for (int k = 0; k < nforl_mpos(LMAX); ++k) AL[k] = f1 * p1[k];
  }
}

template <int LMAX, bool FAST = false>
__device__ __forceinline__ void derivatives(const float (&c)[nforl(LMAX)], float sel, float cel, float saz1, float caz1,
                                            float (&out)[6], float el = 0.0f, const float* tab = nullptr) {
  const double twoc = 2.0 * (double)caz1;
  double cprev = kSqrt2, sprev = 0.0;                                   // m = 0
  double cm = kSqrt2 * (double)caz1, sm = kSqrt2 * (double)saz1;        // m = 1
  const bool atpole = (double)sel < 1e-4;
  float AL[nforl_mpos(LMAX)];
  if constexpr (FAST)
    fast_al<LMAX>(AL, tab, el);
  else
    plm_all<LMAX>(AL, cel, sel * sel, std::make_integer_sequence<int, LMAX + 1>{});
  float dSH_del = 0.0f, dSH_daz = 0.0f, d2SH_del2 = 0.0f, d2SH_deldaz = 0.0f, d2SH_daz2 = 0.0f;
  float amplitude = c[0] * AL[0];
#pragma unroll
  for (int l = 2; l <= LMAX; l += 2) {
    const float v = c[sh_index(l, 0)];
    amplitude += v * AL[index_mpos(l, 0)];
    dSH_del += v * sqrtf((float)(l * (l + 1))) * AL[index_mpos(l, 1)];
    const float t = v * (sqrtf((float)(l * (l + 1) * (l - 1) * (l + 2))) * AL[index_mpos(l, 2)] -
                         (float)(l * (l + 1)) * AL[index_mpos(l, 0)]);
    d2SH_del2 = (float)((double)d2SH_del2 + (double)t / 2.0);
  }
#pragma unroll
  for (int m = 1; m <= LMAX; m++) {
    if (m > 1) {
      const double cn = __builtin_fma(twoc, cm, -cprev), sn = __builtin_fma(twoc, sm, -sprev);
      cprev = cm;
      sprev = sm;
      cm = cn;
      sm = sn;
    }
    const float caz = (float)cm;
    const float saz = (float)sm;
#pragma unroll
    for (int l = (m & 1) ? m + 1 : m; l <= LMAX; l += 2) {
      const float vp = c[sh_index(l, m)];
      const float vm = c[sh_index(l, -m)];
      const float al = AL[index_mpos(l, m)];
      amplitude += (vp * caz + vm * saz) * al;
      float tmp = sqrtf((float)((l + m) * (l - m + 1))) * AL[index_mpos(l, m - 1)];
      if (l > m) tmp -= sqrtf((float)((l - m) * (l + m + 1))) * AL[index_mpos(l, m + 1)];
      tmp = (float)((double)tmp / -2.0);
      dSH_del += (vp * caz + vm * saz) * tmp;
      float tmp2 = (float)(-((l + m) * (l - m + 1) + (l - m) * (l + m + 1))) * al;
      if (m == 1)
        tmp2 -= sqrtf((float)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[index_mpos(l, 1)];
      else
        tmp2 += sqrtf((float)((l + m) * (l - m + 1) * (l + m - 1) * (l - m + 2))) * AL[index_mpos(l, m >= 2 ? m - 2 : 0)];
      if (l > m + 1)
        tmp2 += sqrtf((float)((l - m) * (l + m + 1) * (l - m - 1) * (l + m + 2))) * AL[index_mpos(l, m + 2 <= l ? m + 2 : l)];
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

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {   // Eigen normalize()
  const float n2 = x * x + y * y + z * z;
  if (n2 > 0.0f) {
    const float s = sqrtf(n2);
    x /= s;
    y /= s;
    z /= s;
  }
}

// One iteration of SH.h:514-559 from (x, y, z): the amplitude at the current direction into amp,
// the direction updated in place; true when the step was below the angle tolerance (converged).
template <int LMAX, bool FAST>
__device__ __forceinline__ bool newton_step(const float (&c)[nforl(LMAX)], float& x, float& y, float& z, float& amp,
                                            const float* tab) {
  // the reference takes az = atan2(y, x), el = atan2(hypot(x, y), z) and their sines and cosines;
  // for the unit direction these are hypot(x, y), z, y / hypot and x / hypot (az = 0 at the pole)
  const float rho = hypotf(x, y);
  const float sel = rho, cel = z;
  const float caz = rho > 0.0f ? x / rho : 1.0f, saz = rho > 0.0f ? y / rho : 0.0f;
  float d[6];
  if constexpr (FAST)
    derivatives<LMAX, true>(c, sel, cel, saz, caz, d, atan2f(rho, z), tab);   // the table's row from el itself
  else
    derivatives<LMAX>(c, sel, cel, saz, caz, d);
  float del = sqrtf(d[1] * d[1] + d[2] * d[2]);
  float daz = 0.0f;
  if (del != 0.0f) {
    daz = d[2] / del;
    del = d[1] / del;
  }
  const float dSH_dt = daz * d[2] + del * d[1];
  const float d2SH_dt2 = (float)((double)(daz * daz * d[5]) + 2.0 * (double)daz * (double)del * (double)d[4] +
                                 (double)(del * del * d[3]));
  float dt = d2SH_dt2 != 0.0f ? (-dSH_dt / d2SH_dt2) : (float)kMaxDirChange;
  if (dt < 0.0f) dt = -dt;
  if ((double)dt > kMaxDirChange) dt = (float)kMaxDirChange;
  del *= dt;
  daz *= dt;
  x += del * caz * cel - daz * saz;
  y += del * saz * cel + daz * caz;
  z -= del * sel;
  normalize3(x, y, z);
  amp = d[0];
  return (double)dt < kAngleTolerance;
}

// (value, index) maximum over the wavefront, lower index on ties; every lane gets the result
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(i, off);
    const bool take = (ov > v) || (ov == v && oi < i);
    v = take ? ov : v;
    i = take ? oi : i;
  }
}

constexpr int kRefill = 16;   // idle lanes that trigger a refill (12-16 best on the brain FODs; 32: 7% slower)
constexpr int kGrab = 256;      // items reserved per atomic: one atomic per ~4 refills
constexpr int kGrabTail = 64;   // near the end of the queue: a shorter drain
// one atomic must cover every idle lane (the lanes past the reservation take consecutive items
// from the new block)
static_assert(kGrabTail >= 64 && kGrab >= kGrabTail, "a reservation block holds at least a wavefront of items");

// Searches of voxels [v0, v0 + nv): item g = (v0 + g / nseeds, seed g % nseeds), result in work[g].
// Up to lmax 8 the kernel is held to 168 registers (three wavefronts per SIMD, a few spills), which
// beats two wavefronts without spills; at lmax 10 two fit, at lmax 12 one.
template <int LMAX, bool FAST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LMAX <= 8 ? 3 : (LMAX == 10 ? 2 : 1)))) void peaks_search_kernel(PeaksLaunch L, size_t v0, size_t nv) {
  constexpr int NSH = nforl(LMAX);
  const int lane = threadIdx.x;
  const size_t nitems = nv * (size_t)L.nseeds;
  float c[NSH];
  float x = 0.0f, y = 0.0f, z = 0.0f;
  int it = 0;
  bool busy = false, exhausted = false;
  size_t item = 0;
  unsigned long long rb = 0, re = 0;   // this wavefront's reserved items [rb, re), uniform
  unsigned long long grab = kGrab;
  const unsigned long long tail_window = 2ull * (unsigned long long)kGrab * (unsigned long long)gridDim.x;
  unsigned long long st_it = 0, st_fail = 0;
  for (;;) {
    // lanes without a search take the next items once at least kRefill lanes are idle (or all
    // are): fewer rounds of coefficient loads for a short idle wait.  Items come from the
    // wavefront's reservation [rb, re), topped up by `grab` items per atomic: kGrab, then kGrabTail
    // once fewer than two rounds of kGrab reservations per wavefront are left, so that the last
    // wavefronts do not hold four searches per lane each while the others have finished.
    const unsigned long long need = __ballot(!busy);
    const unsigned nneed = (unsigned)__popcll(need);
    if (!exhausted && (nneed >= (unsigned)kRefill || need == ~0ull)) {
      const unsigned long long rem = re - rb;
      unsigned long long base = 0;
      if (rem < nneed) {
        const int first = __ffsll((long long)need) - 1;
        if (lane == first) base = atomicAdd(L.counter, grab);
        base = readlane64(base, first);
        if (base >= nitems) exhausted = true;
      }
      if (!busy) {
        const unsigned long long below = lane ? (need & (~0ull >> (64 - lane))) : 0ull;
        const unsigned long long rank = (unsigned long long)__popcll(below);
        const unsigned long long pick = rank < rem ? rb + rank : base + (rank - rem);
        item = (size_t)(pick < nitems ? pick : nitems);
        if (item < nitems) {
          const size_t vox = v0 + item / (size_t)L.nseeds;
          const int s = (int)(item % (size_t)L.nseeds);
          const float* src = L.sh + vox * (size_t)L.ld_sh;
          bool bad = L.mask && !L.mask[vox];
          bool nonzero = false;
#pragma unroll
          for (int j = 0; j < NSH; ++j) {
            c[j] = src[j];
            bad = bad || isnan(c[j]);
            nonzero = nonzero || (j > 0 && c[j] != 0.0f);
          }
          if (L.match && L.ld_match && isnan(L.match[vox * (size_t)L.ld_match])) bad = true;
          if (bad || !nonzero) {
            L.work[item] = make_float4(NAN, NAN, NAN, NAN);   // the voxel is reported as NaN
          } else {
            x = L.seeds[3 * s];
            y = L.seeds[3 * s + 1];
            z = L.seeds[3 * s + 2];
            it = 0;
            busy = true;
          }
        }
      }
      // the reservation left (re <= nitems, so every item taken from it exists; an atomic past
      // the end leaves it empty and the wavefront exhausted)
      if (rem < nneed) {
        const unsigned long long hi = base + grab < nitems ? base + grab : nitems;
        const unsigned long long lo = base + (nneed - rem);
        rb = lo < hi ? lo : hi;
        re = hi;
        if (base + tail_window >= nitems) grab = kGrabTail;
      } else {
        rb += nneed;
      }
    }
    if (__ballot(busy) == 0ull) {
      if (exhausted) break;
      continue;
    }
    if (busy) {
      float amp;
      const bool conv = newton_step<LMAX, FAST>(c, x, y, z, amp, L.fast_table);
      ++it;
      if (conv || it == 50) {
        L.work[item] = conv ? make_float4(amp, x, y, z) : make_float4(NAN, NAN, NAN, NAN);
        st_it += (unsigned long long)it;
        if (!conv) ++st_fail;
        busy = false;
      }
    }
  }
  if (L.stats) {
    const unsigned long long itsum = wave_sum(st_it), fail = wave_sum(st_fail);
    if (lane == 0) {
      atomicAdd(L.stats + 1, itsum);
      atomicAdd(L.stats + 2, fail);
    }
  }
}

// Accepted and reported peaks of voxels [v0, v0 + nv) from their searches' results.
__global__ __launch_bounds__(64) void peaks_select_kernel(PeaksLaunch L, size_t v0, size_t nv) {
  __shared__ float4 s_cand[64];                 // the batch's candidates (amplitude, direction)
  __shared__ float4 s_acc[kPeaksMaxAccepted];   // accepted peaks, in seed order
  const int lane = threadIdx.x;
  const int nout_vals = 3 * L.npeaks;
  unsigned long long st_vox = 0, st_over = 0;

  // voxels are taken in a fixed stride (the work per voxel is nearly uniform; a shared counter
  // would serialise one atomic per voxel on one address)
  for (size_t t = blockIdx.x; t < nv; t += gridDim.x) {
    const size_t vox = v0 + t;
    float* o = L.out + vox * (size_t)nout_vals;

    // ---- input checks: mask (DataLoader, sh2peaks.cpp:126-130) and check_input (:266-286) ----
    const float* src = L.sh + vox * (size_t)L.ld_sh;
    const float v0c = lane < L.n_sh ? src[lane] : 0.0f;
    const float v1c = lane + 64 < L.n_sh ? src[lane + 64] : 0.0f;
    const bool has_nan = __ballot(isnan(v0c) || isnan(v1c)) != 0ull;
    const bool nonzero = __ballot((lane >= 1 && v0c != 0.0f) || v1c != 0.0f) != 0ull;
    const float* match = L.match ? L.match + (L.ld_match ? vox * (size_t)L.ld_match : 0) : nullptr;
    const bool masked = L.mask && !L.mask[vox];
    const bool match_nan = match && L.ld_match && isnan(match[0]);
    if (masked || has_nan || !nonzero || match_nan) {
      for (int k = lane; k < nout_vals; k += 64) o[k] = NAN;
      continue;
    }
    ++st_vox;

    // ---- the searches' results, accepted in seed order (sh2peaks.cpp:176-189) ----
    // Lane j holds candidate j of a batch of 64 seeds.  For candidate i in seed order, one ballot
    // gives the candidates within acos(0.99) of it (|dot| is symmetric: the same float value
    // either way round), so it is a duplicate when that mask meets the accepted set A; validity,
    // the threshold and the earlier batches' accepted peaks are per-lane tests done up front.
    // The dot product is the reference's p.dot(peaks[i]) term for term, and (double)|dot| > 0.99
    // is |dot| >= 0.99f (the float nearest 0.99 lies above it).
    const float4* res = L.work + (size_t)t * (size_t)L.nseeds;
    int nacc = 0, nstored = 0;
    for (int b0 = 0; b0 < L.nseeds; b0 += 64) {
      const int s = b0 + lane;
      const float4 r = s < L.nseeds ? res[s] : make_float4(NAN, NAN, NAN, NAN);
      const int nb = L.nseeds - b0 < 64 ? L.nseeds - b0 : 64;
      bool dup_prev = false;   // within acos(0.99) of a peak accepted in an earlier batch
      for (int k = 0; k < nstored; ++k) {
        const float4 q = s_acc[k];
        dup_prev = dup_prev || fabsf(r.y * q.y + r.z * q.z + r.w * q.w) >= 0.99f;
      }
      // finite, not within acos(0.99) of an earlier batch's peak, and at or above the threshold:
      // the candidates that are accepted unless within acos(0.99) of one accepted in this batch
      const unsigned long long cand = __ballot(isfinite(r.x) && !dup_prev && r.x >= L.threshold);
      s_cand[lane] = r;   // every lane writes: entries past nb hold NaN candidates
      wave_sync();
      const int room = kPeaksMaxAccepted - nstored;
      unsigned long long A = 0ull;   // accepted and stored in this batch
      if (nb <= room) {
        // every peak accepted in this batch is stored: the decision per candidate is one mask
        // test (candidates past nb have no bit in cand)
        for (int i0 = 0; i0 < nb; i0 += 4) {   // i0 + 3 <= 63: four LDS reads issued together
          float4 q[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) q[u] = s_cand[i0 + u];
          unsigned long long close[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) close[u] = __ballot(fabsf(r.y * q[u].y + r.z * q[u].z + r.w * q[u].w) >= 0.99f);
#pragma unroll
          for (int u = 0; u < 4; ++u) A |= (close[u] & A) ? 0ull : (cand & (1ull << (i0 + u)));
        }
        nacc += __popcll(A);
      } else {
        for (int i = 0; i < nb; ++i) {
          const float4 q = s_cand[i];
          const unsigned long long close = __ballot(fabsf(r.y * q.y + r.z * q.z + r.w * q.w) >= 0.99f);
          if (((cand >> i) & 1ull) && !(close & A)) {
            if (__popcll(A) < room) A |= 1ull << i;
            ++nacc;
          }
        }
      }
      if ((A >> lane) & 1ull) {
        const unsigned long long below = lane ? (A & (~0ull >> (64 - lane))) : 0ull;
        s_acc[nstored + __popcll(below)] = r;
      }
      nstored += __popcll(A);
      wave_sync();
    }
    if (nacc > kPeaksMaxAccepted) {
      ++st_over;
      nacc = kPeaksMaxAccepted;
    }
    float my_a = NAN, my_x = 0.0f, my_y = 0.0f, my_z = 0.0f;   // accepted peak `lane`
    if (lane < nstored) {
      const float4 q = s_acc[lane];
      my_a = q.x;
      my_x = q.y;
      my_y = q.z;
      my_z = q.w;
    }
    wave_sync();

    // ---- the reported peaks (sh2peaks.cpp:204-246) ----
    const int nrep = L.npeaks < nacc ? L.npeaks : nacc;
    unsigned long long used = 0ull;
    for (int i = 0; i < nrep; ++i) {
      float key;
      if (match) {
        // the accepted peak most nearly parallel to the given direction (|dot| > 0, first on ties)
        float px = match[3 * i], py = match[3 * i + 1], pz = match[3 * i + 2];
        normalize3(px, py, pz);
        key = lane < nacc ? fabsf(px * my_x + py * my_y + pz * my_z) : -1.0f;
      } else {
        // decreasing amplitude (partial_sort_copy), first accepted on ties
        key = (lane < nacc && !((used >> lane) & 1ull)) ? my_a : -INFINITY;
      }
      int idx = lane < nacc ? lane : INT_MAX;
      wave_argmax(key, idx);
      const int sel = uniform32(idx);
      const bool ok = sel < nacc && (match ? key > 0.0f : true);
      if (ok) used |= 1ull << sel;
      const float a = ok ? fbits(my_a, sel) : NAN;
      const float px = ok ? fbits(my_x, sel) : NAN, py = ok ? fbits(my_y, sel) : NAN, pz = ok ? fbits(my_z, sel) : NAN;
      if (lane == 0) o[3 * i] = a * px;
      if (lane == 1) o[3 * i + 1] = a * py;
      if (lane == 2) o[3 * i + 2] = a * pz;
    }
    for (int k = 3 * nrep + lane; k < nout_vals; k += 64) o[k] = NAN;
  }
  if (L.stats && lane == 0) {
    atomicAdd(L.stats + 0, st_vox);
    atomicAdd(L.stats + 3, st_over);
  }
}

__global__ void reset_counter_kernel(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

template <class K>
hipError_t persistent_grid(K kern, size_t work, size_t* grid) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  size_t g = (size_t)ncu * (size_t)per_cu;
  *grid = g < work ? g : (work ? work : 1);
  return hipSuccess;
}

template <int LMAX>
hipError_t launch_lmax(const PeaksLaunch& L, hipStream_t stream) {
  const size_t chunk = L.work_items / (size_t)L.nseeds;
  for (size_t v0 = 0; v0 < L.nvox; v0 += chunk) {
    const size_t nv = L.nvox - v0 < chunk ? L.nvox - v0 : chunk;
    size_t g1 = 0, g2 = 0;
    hipError_t e = L.fast_table ? persistent_grid(peaks_search_kernel<LMAX, true>, nv * (size_t)L.nseeds / 64 + 1, &g1)
                                : persistent_grid(peaks_search_kernel<LMAX, false>, nv * (size_t)L.nseeds / 64 + 1, &g1);
    if (e != hipSuccess) return e;
    if ((e = persistent_grid(peaks_select_kernel, nv, &g2)) != hipSuccess) return e;
    reset_counter_kernel<<<1, 64, 0, stream>>>(L.counter);
    if (L.fast_table)
      peaks_search_kernel<LMAX, true><<<dim3((unsigned)g1), dim3(64), 0, stream>>>(L, v0, nv);
    else
      peaks_search_kernel<LMAX, false><<<dim3((unsigned)g1), dim3(64), 0, stream>>>(L, v0, nv);
    peaks_select_kernel<<<dim3((unsigned)g2), dim3(64), 0, stream>>>(L, v0, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t peaks_launch(const PeaksLaunch& L, hipStream_t stream) {
  if (L.n_sh < nforl(L.lmax) || L.n_sh > 128 || L.nseeds < 1 || L.nseeds > kPeaksMaxSeeds || L.npeaks < 1 ||
      L.npeaks > kPeaksMaxOut || !L.work || L.work_items < (size_t)L.nseeds)
    return hipErrorInvalidValue;
  switch (L.lmax) {
    case 2: return launch_lmax<2>(L, stream);
    case 4: return launch_lmax<4>(L, stream);
    case 6: return launch_lmax<6>(L, stream);
    case 8: return launch_lmax<8>(L, stream);
    case 10: return launch_lmax<10>(L, stream);
    case 12: return launch_lmax<12>(L, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mrcsd
