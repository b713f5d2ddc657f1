// Batched inequality-constrained least squares on gfx950: Math::ICLS::Solver::operator()
// (core/math/constrained_least_squares.h:237-357) for one right-hand side per voxel.
//
// One wavefront solves one voxel at a time and pulls the next from a work counter (persistent).
// The dual active-set iteration is the reference's, step for step: add the most violated
// inequality (lowest index on ties), solve (B_act B_act^T + lambda I) l = -c_u[act], drop the
// active inequality with the smallest lambda_prev / (lambda_prev - l) (lowest index on ties) while
// any multiplier is negative, stop when the added constraint was already active or after
// max_niter outer iterations.  What differs is how each step is computed:
//   * active constraints live in slots, one per lane; B_act B_act^T is gathered from the
//     precomputed Gram matrix G = B B^T instead of being re-formed from B (k^2 n flops saved);
//   * the k x k system is solved by a Gauss-Jordan sweep, thread a owning row a;
//   * constraint values are updated as c = c_u + G[:, act] l, which equals B x - t for
//     x = y_u + B_act^T l;
//   * the final x = L^-T (y_u + B_act^T l) is formed as the reference forms it: y gathered from
//     B^T, then a back substitution with L^T, one row per step (a precomputed L^-1 would lose
//     cond(L)^2 rather than cond(L) of accuracy on nearly singular problems).
// All arithmetic is FP64; results agree with the reference to rounding.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "icls_solve.h"
#include "wave_utils.h"

namespace mrcsd {
namespace {

constexpr int KMAX = kIclsMaxActive;
constexpr int kQ = kIclsMaxCon / 64;

__device__ __forceinline__ bool test_bit(const unsigned long long* w, int i) {
  unsigned long long word = 0ull;
#pragma unroll
  for (int q = 0; q < kQ; ++q) word = (q == (i >> 6)) ? w[q] : word;
  return (word >> (i & 63)) & 1ull;
}

__device__ __forceinline__ void set_bit(unsigned long long* w, int i, bool on) {
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const unsigned long long m = 1ull << (i & 63);
    if (q == (i >> 6)) w[q] = on ? (w[q] | m) : (w[q] & ~m);
  }
}

__global__ __launch_bounds__(64) void icls_persistent_kernel(IclsDeviceProblem P, IclsLaunch L) {
  __shared__ double s_b[kIclsMaxMeas];
  __shared__ double s_y[64];
  __shared__ double s_cu[kIclsMaxCon];   // c_u, indexed by constraint
  __shared__ double s_l[KMAX];
  __shared__ int s_act[KMAX];
  __shared__ double s_col[2][64];
  __shared__ unsigned long long s_st[8];

  const int tid = threadIdx.x;
  const int n = P.n, nm = P.nmeas, ncon = P.ncon, ldg = P.ncon_pad;
  const int nq = P.ncon_pad / 64;   // host guarantees ncon_pad % 64 == 0, ncon_pad <= kIclsMaxCon
  const int num_ineq = ncon - P.num_eq;
  if (tid < 8) s_st[tid] = 0ull;

  for (;;) {
    unsigned long long vox = 0;
    if (tid == 0) vox = atomicAdd(L.counter, 1ull);
    vox = uniform64(vox);
    if (vox >= L.nvox) break;
    if (L.valid && !L.valid[vox]) {
      if (tid < n) L.x[vox * n + tid] = 0.0f;
      if (tid == 0) {
        if (L.iters) L.iters[vox] = 0xFFFF;
        if (L.nactive) L.nactive[vox] = 0;
      }
      continue;
    }
    const float* bsrc = L.b + vox * (size_t)L.ld_b;
    for (int r = tid; r < nm; r += 64) s_b[r] = (double)bsrc[r];
    wave_sync();

    // ---- unconstrained solution and its constraint values (:244-248) ----
    double yu = 0.0;
    if (tid < n)
      for (int r = 0; r < nm; ++r) yu = fma(P.b2d[r * n + tid], s_b[r], yu);
    s_y[tid] = yu;
    wave_sync();
    double c[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      c[q] = 0.0;
      if (q < nq) {
        const int r = tid + 64 * q;
        const double* bp = P.Bt + r;
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc = fma(bp[(size_t)j * ldg], s_y[j], acc);
        acc -= (r < ncon) ? P.t[r] : 0.0;
        c[q] = acc;
        s_cu[r] = acc;   // s_cu holds ncon_pad entries
      }
    }

    // ---- active set: equality constraints start (and stay) active (:259-261) ----
    int k = P.num_eq;
    int my_act = tid < k ? num_ineq + tid : -1;
    double lamprev = 0.0, la = 0.0;
    if (tid < k) s_act[tid] = my_act;
    unsigned long long actw[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) actw[q] = 0ull;
    for (int e = 0; e < P.num_eq; ++e) set_bit(actw, num_ineq + e, true);
    int niter = 0;
    bool fail = false;
    wave_sync();

    for (;;) {
      // most violated inequality, lowest index on ties (:270)
      double bv = INFINITY;
      int bi = INT_MAX;
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        const int r = tid + 64 * q;
        if (q < nq && r < num_ineq && c[q] < bv) { bv = c[q]; bi = r; }
      }
      wave_argmin(bv, bi);
      bi = uniform32(bi);
      if (!(bv < -P.tol)) break;
      bool changed = !test_bit(actw, bi);
      if (changed) {
        if (k >= KMAX) { fail = true; break; }
        if (tid == k) { my_act = bi; lamprev = 0.0; s_act[k] = bi; }
        ++k;
        set_bit(actw, bi, true);
      }

      for (;;) {   // inner loop (:274-331)
        wave_sync();
        const bool slot = tid < k;
        // row `tid` of B_act B_act^T + lambda I, gathered from G
        const double* grow = P.G + (size_t)(slot ? my_act : 0) * ldg;
        double A[KMAX];
#pragma unroll
        for (int bb = 0; bb < KMAX; ++bb) {
          double v = 0.0;
          if (bb < k) {
            const double g = grow[s_act[bb]];
            v = slot ? g : 0.0;
          }
          A[bb] = (bb == tid) ? v + P.lambda_min : v;
        }
        double rhs = slot ? -s_cu[my_act] : 0.0;
        // Gauss-Jordan sweep (thread a owns row a); steps past k are skipped by uniform branches
#pragma unroll
        for (int kk = 0; kk < KMAX; ++kk) {
          if (kk < k) {
            const int buf = kk & 1;
            s_col[buf][tid] = A[kk];
            const double d = lane_bcast(A[kk], kk);
            const double bk = lane_bcast(rhs, kk);
            wave_sync();
            const double cf = (tid != kk) ? A[kk] / d : 0.0;
#pragma unroll
            for (int j = kk + 1; j < KMAX; ++j) A[j] = fma(-cf, s_col[buf][j], A[j]);
            rhs = fma(-cf, bk, rhs);
          }
        }
        double di = 1.0;
#pragma unroll
        for (int j = 0; j < KMAX; ++j) di = (tid == j) ? A[j] : di;
        la = slot ? rhs / di : 0.0;
        if (tid == 0) {
          const unsigned long long kk = (unsigned long long)k;
          s_st[2] += 1ull;
          s_st[3] += kk;
          s_st[4] += kk * kk;
          s_st[5] += kk * kk * kk;
        }

        // worst offender among negative inequality multipliers (:297-313)
        double sv = INFINITY;
        int si = INT_MAX;
        if (slot && my_act < num_ineq && la < 0.0) {
          sv = lamprev / (lamprev - la);
          si = my_act;
        }
        wave_argmin(sv, si);
        if (!(sv < INFINITY)) break;   // all feasible: x = y_u + B_act^T l (:316-321)
        si = uniform32(si);
        changed = true;                // si is active by construction (:327-329)
        set_bit(actw, si, false);
        const int pos = uniform32(__ffsll((long long)__ballot(slot && my_act == si)) - 1);
        const int last = k - 1;
        const int act_last = lane_bcast(my_act, last);
        const double lp_last = lane_bcast(lamprev, last);
        if (tid == pos) { my_act = act_last; lamprev = lp_last; }
        if (tid == last) { my_act = -1; lamprev = 0.0; }
        --k;
        wave_sync();
        if (tid < k) s_act[tid] = my_act;
      }
      if (fail) break;

      lamprev = la;          // lambda_prev = lambda (:334)
      ++niter;
      if (!changed || niter > P.max_niter) break;

      // c = B x - t = c_u + G[:, act] l (:349-351)
      wave_sync();
      s_l[tid] = la;
      if (tid < k) s_act[tid] = my_act;
      wave_sync();
      if (tid == 0) s_st[6] += (unsigned long long)k;
#pragma unroll
      for (int q = 0; q < kQ; ++q) c[q] = (q < nq) ? s_cu[tid + 64 * q] : 0.0;
      for (int a = 0; a < k; ++a) {
        const double lv = s_l[a];
        const double* g = P.G + (size_t)s_act[a] * ldg + tid;
#pragma unroll
        for (int q = 0; q < kQ; ++q)
          if (q < nq) c[q] = fma(g[64 * q], lv, c[q]);
      }
    }

    // ---- x = y_u + B_act^T l (:319), then L^T x = y solved in place (:355) ----
    wave_sync();
    s_l[tid] = tid < k ? la : 0.0;
    s_act[tid] = tid < k ? my_act : 0;
    wave_sync();
    double yv = 0.0;   // lane j < n: y_j, then x_j
    if (tid < n) {
      yv = s_y[tid];
      for (int a = 0; a < k; ++a) yv = fma(P.Bt[(size_t)tid * ldg + s_act[a]], s_l[a], yv);
    }
    // back substitution by columns: step i finalises x_i and removes it from the rows above
    for (int i = n - 1; i >= 0; --i) {
      const double xi = lane_bcast(yv, i) / P.L[(size_t)i * n + i];
      if (tid == i) yv = xi;
      if (tid < i) yv = fma(-P.L[(size_t)i * n + tid], xi, yv);
    }
    if (tid < n) L.x[vox * n + tid] = (float)yv;
    if (tid == 0) {
      if (L.iters) L.iters[vox] = fail ? (uint16_t)0xFFFE : (uint16_t)niter;
      if (L.nactive) L.nactive[vox] = (uint16_t)k;
    }
    if (tid == 0) {
      s_st[0] += 1ull;
      s_st[1] += (unsigned long long)niter;
      s_st[7] += fail ? 1ull : 0ull;
    }
    wave_sync();
  }
  wave_sync();
  if (L.stats && tid < 8) atomicAdd(L.stats + tid, s_st[tid]);
}

__global__ void reset_counter_kernel(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

}  // namespace

hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, hipStream_t stream) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, icls_persistent_kernel, 64, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  size_t grid = (size_t)ncu * (size_t)per_cu;
  if (grid > L.nvox) grid = L.nvox ? L.nvox : 1;
  reset_counter_kernel<<<1, 64, 0, stream>>>(L.counter);
  icls_persistent_kernel<<<dim3((unsigned)grid), dim3(64), 0, stream>>>(P, L);
  return hipGetLastError();
}

}  // namespace mrcsd
