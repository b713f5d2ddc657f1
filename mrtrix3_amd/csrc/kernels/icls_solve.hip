// Batched inequality-constrained least squares on gfx950: Math::ICLS::Solver::operator()
// (core/math/constrained_least_squares.h:237-357) for one right-hand side per voxel.
//
// One wavefront solves one voxel at a time and pulls the next from a work counter (persistent).
// The dual active-set iteration is the reference's, step for step: add the most violated
// inequality (lowest index on ties), solve (B_act B_act^T + lambda I) l = -c_u[act], drop the
// active inequality with the smallest lambda_prev / (lambda_prev - l) (lowest index on ties) while
// any multiplier is negative, stop when the added constraint was already active or after
// max_niter outer iterations.  What differs is how each step is computed:
//
//   * The active set lives in slots, slot s on lane s % 64 (register s / 64), in the order the
//     constraints entered.  The Cholesky factor L of B_act B_act^T + lambda I is kept from one
//     inner solve to the next instead of being re-factorised (the reference forms the matrix and
//     its LLT afresh every time, k^2 n + k^3 / 3 flops): row s of L is a row of LDS owned by slot s
//     (an indirection, so that deleting a slot moves no rows).
//       - adding constraint j: L w = G[act, j] by forward substitution, new row [w, sqrt(G_jj +
//         lambda - w.w)];
//       - removing slot p: delete row p, then Givens rotations on column pairs (c, c+1),
//         c = p..k-2, restore the triangle (the rotated-out last column is dropped);
//       - z = L^-1 (-c_u[act]) is updated alongside (one more entry on an addition; the same
//         rotations on a removal: z behaves as one more row of L), so each solve is one backward
//         substitution L^T l = z.
//     Both substitutions run one column per step: the lane that owns the pivot finishes its
//     entry, a readlane hands it to the wavefront, and every lane updates its own entry with one
//     fused multiply-add (the coefficient, with the pivot's reciprocal folded in, is computed off
//     that chain).
//   * Rows of B_act B_act^T come from the precomputed Gram matrix G = B B^T, and the constraint
//     values as c = c_u + G[:, act] l, which equals B x - t for x = y_u + B_act^T l.
//   * A re-solve with an unchanged active set (the most violated constraint already active) is
//     the previous solve, whose multipliers are all non-negative; the loop stops there, as the
//     reference does after repeating it.
//   * The final x = L_H^-T (y_u + B_act^T l) is formed as the reference forms it: y gathered from
//     B^T, then a back substitution with L_H^T, one row per step.
// All arithmetic is FP64; results agree with the reference to rounding.
//
// Capacity: a pass holds at most KCAP active constraints (its LDS factor is KCAP rows).  A voxel
// that needs more is queued for the next pass, which solves it again from the start with more
// slots (the arithmetic does not depend on the capacity); after the largest capacity (128) a voxel
// is reported with iters 0xFFFE.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "icls_solve.h"
#include "wave_utils.h"

namespace mrcsd {
namespace {

struct IclsPass {
  const uint32_t* in_list;             // voxels of this pass (null: every voxel, 0..nvox-1)
  const unsigned long long* in_count;  // their number (device)
  uint32_t* out_list;                  // voxels that outgrow this pass (null: last pass)
  unsigned long long* out_count;
  unsigned long long* counter;         // work counter of this pass
};

// Slot s's value of a per-slot register array (s wave-uniform).
template <int SPL>
__device__ __forceinline__ double slot_get(const double (&a)[SPL], int s) {
  double v = a[0];
#pragma unroll
  for (int j = 1; j < SPL; ++j) v = ((s >> 6) == j) ? a[j] : v;
  return lane_bcast(v, s & 63);
}

template <int SPL>
__device__ __forceinline__ int slot_get(const int (&a)[SPL], int s) {
  int v = a[0];
#pragma unroll
  for (int j = 1; j < SPL; ++j) v = ((s >> 6) == j) ? a[j] : v;
  return lane_bcast(v, s & 63);
}

// Slots p..last-1 take the value of the next slot.
template <int SPL, class T>
__device__ __forceinline__ void shift_down(T (&a)[SPL], int lane, int p, int last) {
  T nxt[SPL];
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    nxt[j] = __shfl(a[j], (lane + 1) & 63);
    const T wrap = (j + 1 < SPL) ? lane_bcast(a[j + 1 < SPL ? j + 1 : j], 0) : nxt[j];
    if (lane == 63) nxt[j] = wrap;
  }
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const int s = lane + 64 * j;
    if (s >= p && s < last) a[j] = nxt[j];
  }
}

__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// 1/sqrt(h): the hardware estimate refined by two Newton steps (relative error near 2^-53).
__device__ __forceinline__ double rsqrt_refined(double h) {
  double y = __builtin_amdgcn_rsq(h);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-h * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// IO64: measurements read from L.b64 and solutions written to L.x64 in double (single precision
// otherwise, the dwi2fod path).
template <int KCAP, int KQ, bool IO64>
__global__ __launch_bounds__(64) void icls_active_set_kernel(IclsDeviceProblem P, IclsLaunch L, IclsPass Q) {
  constexpr int SPL = (KCAP + 63) / 64;           // slots per lane
  __shared__ double s_W[KCAP * (KCAP + 1) / 2];   // inverse factor W = L^-1, packed lower rows
  // per-slot values staged for broadcast reads: a vector (s_v[s]), rotation parameters
  // (s_v[2c], s_v[2c + 1]) or (multiplier, constraint index) pairs; with the factor, under 10 KB
  // at 48 slots, so that four wavefronts share each SIMD
  __shared__ __attribute__((aligned(16))) double s_v[2 * KCAP];

  const int lane = threadIdx.x;
  const int n = P.n, nm = P.nmeas, ldg = P.ncon_pad;
  // constraint r on lane r % 64, word r / 64: ncon_pad = 64 KQ (host), padding rows zero
  const int num_ineq = P.ncon - P.num_eq;
  const double lam = P.lambda_min;

  // statistics of this wavefront (wave-uniform)
  unsigned long long st_vox = 0, st_outer = 0, st_inner = 0, st_k = 0, st_k2 = 0, st_k3 = 0, st_kc = 0,
                     st_ovf = 0;

  // start of packed row r
  auto T = [](int r) __attribute__((always_inline)) { return r * (r + 1) / 2; };

  for (;;) {
    unsigned long long idx = 0;
    if (lane == 0) idx = atomicAdd(Q.counter, 1ull);
    idx = uniform64(idx);
    size_t vox;
    if (Q.in_list) {
      if (idx >= *Q.in_count) break;
      vox = (size_t)uniform32((int)Q.in_list[idx]);
    } else {
      if (idx >= L.nvox) break;
      vox = (size_t)idx;
      if (L.valid && !L.valid[vox]) {
        for (int j = lane; j < n; j += 64) {
          if constexpr (IO64) L.x64[vox * n + j] = 0.0;
          else L.x[vox * n + j] = 0.0f;
        }
        if (lane == 0) {
          if (L.iters) L.iters[vox] = 0xFFFF;
          if (L.nactive) L.nactive[vox] = 0;
        }
        continue;
      }
    }
    const auto* bsrc = [&]() {
      if constexpr (IO64) return L.b64 + vox * (size_t)L.ld_b;
      else return L.b + vox * (size_t)L.ld_b;
    }();

    // ---- unconstrained solution y_u = b2d^T b (:244), parameter j on lane j % 64 ----
    double yu[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = lane + 64 * h;
      double acc = 0.0;
      if (h == 0 || n > 64) {
        const int jj = j < n ? j : 0;
#pragma unroll 8
        for (int r = 0; r < nm; ++r) acc = fma(P.b2d[(size_t)r * n + jj], (double)bsrc[r], acc);
      }
      yu[h] = j < n ? acc : 0.0;
    }
    // ---- constraint values of y_u: c_u = B y_u - t (:246-248), constraint lane + 64 q ----
    double cu[KQ];
#pragma unroll
    for (int q = 0; q < KQ; ++q) cu[q] = 0.0;
    for (int j = 0; j < n; ++j) {
      const double yj = slot_get<2>(yu, j);
      const double* bt = P.Bt + (size_t)j * ldg + lane;
#pragma unroll
      for (int q = 0; q < KQ; ++q)
        cu[q] = fma(bt[64 * q], yj, cu[q]);
    }
#pragma unroll
    for (int q = 0; q < KQ; ++q)
      cu[q] -= P.t[lane + 64 * q];

    // ---- active set: empty, equality constraints active from the start (:255-261) ----
    int k = 0;
    int aidx[SPL];
    double lp[SPL], lv[SPL];   // lambda_prev and the current multipliers l = A^-1 (-c_u[act]) of each slot
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
      aidx[j] = -1;
      lp[j] = lv[j] = 0.0;
    }
    unsigned long long actw = 0ull;   // lane w: bit i of constraint 64 w + i

    // Stages x[slot] for slots < kk in s_v.
    auto stage = [&](const double (&x)[SPL], int kk) __attribute__((always_inline)) {
      wave_sync();   // earlier readers of s_v are done
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj)
        if (lane + 64 * jj < kk) s_v[lane + 64 * jj] = x[jj];
      wave_sync();
    };
    // (multiplier, constraint index) of each active slot in s_v
    auto stage_pairs = [&]() __attribute__((always_inline)) {
      wave_sync();
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int s = lane + 64 * jj;
        if (s < k) {
          s_v[2 * s] = lv[jj];
          s_v[2 * s + 1] = (double)aidx[jj];
        }
      }
      wave_sync();
    };
    // out_i = sum_{m <= i} W[i][m] s_v[m] for rows i < kk (a row of W per slot lane; lanes past
    // kk give 0).
    auto mv_rows = [&](double (&out)[SPL], int kk) __attribute__((always_inline)) {
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int i = lane + 64 * jj;
        const double* row = s_W + (i < kk ? T(i) : 0);
        double acc = 0.0;
        if (i < kk) {
#pragma unroll 4
          for (int m = 0; m <= i; ++m) acc = fma(row[m], s_v[m], acc);
        }
        out[jj] = acc;
      }
    };
    // out_c = sum_{m >= c} W[m][c] s_v[m] for columns c < kk (W^T x; rows read contiguously;
    // lanes past kk give 0).
    auto mv_cols = [&](double (&out)[SPL], int kk) __attribute__((always_inline)) {
      double acc[SPL];
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) acc[jj] = 0.0;
      int tm = 0;   // start of row m
#pragma unroll 4
      for (int m = 0; m < kk; ++m) {
        const double xm = s_v[m];
#pragma unroll
        for (int jj = 0; jj < SPL; ++jj)
          if (lane + 64 * jj <= m) acc[jj] = fma(s_W[tm + lane + 64 * jj], xm, acc[jj]);
        tm += m + 1;
      }
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) out[jj] = acc[jj];
    };

    // Appends constraint j as slot k.  With L w = g (g = G[act, j]), the new factor row is
    // [w, d], d = sqrt(G_jj + lambda - w.w); the inverse factor gains the row u = [-(w^T W) / d,
    // 1 / d].  With z = W rhs (never formed) the new entry of z is (rhs_j - w.z) / d, where
    // w.z = g^T A^-1 rhs = g.l, and l = W^T z changes by u^T z_k.
    auto append = [&](int j) __attribute__((always_inline)) {
      const double* gj = P.G + (size_t)j * ldg;
      double g[SPL], w[SPL], nr[SPL];
      double gl = 0.0;
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const double v = gj[aidx[jj] >= 0 ? aidx[jj] : 0];
        g[jj] = (lane + 64 * jj < k) ? v : 0.0;
        gl = fma(g[jj], lv[jj], gl);
      }
      const double gjj = gj[j] + lam;
      stage(g, k);
      mv_rows(w, k);                       // w = W g
      double ww = 0.0;
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) ww = fma(w[jj], w[jj], ww);
      ww = wave_sum_d(ww);
      gl = wave_sum_d(gl);
      const double y = rsqrt_refined(gjj - ww);   // 1 / d
      // right-hand side -c_u[j] (each word read out by readlane and the wanted one selected in
      // scalar registers: a select over the lane's words would be lowered to an indexed scratch array)
      double cj = 0.0;
#pragma unroll
      for (int q = 0; q < KQ; ++q) {
        const double v = lane_bcast(cu[q], j & 63);
        cj = ((j >> 6) == q) ? v : cj;
      }
      const double zk = (-cj - gl) * y;
      stage(w, k);
      mv_cols(nr, k);                      // w^T W
      const int base = T(k);
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int c = lane + 64 * jj;
        const double u = -y * nr[jj];
        if (c < k) {
          s_W[base + c] = u;
          lv[jj] = fma(u, zk, lv[jj]);
        }
        if (c == k) {
          s_W[base + c] = y;
          aidx[jj] = j;
          lp[jj] = 0.0;
          lv[jj] = y * zk;
        }
      }
      ++k;
    };

    // Removes slot p.  Deleting row p of L and rotating column pairs (c, c+1), c = p..k-2, back to a
    // triangle is, for W, moving column p to the end and rotating rows (c, c+1) so that the moved
    // column vanishes: the rotation of step c takes a_c (the carried row's last entry; W[p][p] at
    // the first step) and b_c = W[c+1][p], and leaves r_c = sqrt(a_c^2 + b_c^2) as the next a.  So
    // r_c^2 is a prefix sum of W[p..c+1][p]^2 and every rotation is known before any is applied.
    // The multipliers of the smaller system are l' = l - (l_p / B_pp) B e_p (B = A^-1 = W^T W,
    // B e_p = W^T W e_p, B_pp = |W e_p|^2), slot p dropped.
    auto remove = [&](int p) __attribute__((always_inline)) {
      const int K = k;
      double wp[SPL], q[SPL], bp[SPL];
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int m = lane + 64 * jj;
        const bool on = m >= p && m < K;
        const double v = s_W[on ? T(m) + p : 0];
        wp[jj] = on ? v : 0.0;
        q[jj] = wp[jj] * wp[jj];
      }
      stage(wp, K);
      mv_cols(bp, K);                      // B e_p
      // inclusive prefix sums of q over the slots
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const double t = __shfl(q[jj], (lane - off) & 63);
          if (lane >= off) q[jj] += t;
        }
      }
#pragma unroll
      for (int jj = 1; jj < SPL; ++jj) q[jj] += lane_bcast(q[jj - 1], 63);
      {
        const double lpv = slot_get<SPL>(lv, p);
        const double bpp = slot_get<SPL>(q, K - 1);   // |W e_p|^2
        const double f = lpv / bpp;
#pragma unroll
        for (int jj = 0; jj < SPL; ++jj) lv[jj] = fma(-f, bp[jj], lv[jj]);
      }
      wave_sync();   // mv_cols has read the staged column
      double qn[SPL], wn[SPL];   // the next slot's values
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        qn[jj] = q[jj];
        wn[jj] = wp[jj];
      }
      shift_down<SPL>(qn, lane, 0, KCAP);
      shift_down<SPL>(wn, lane, 0, KCAP);
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int c = lane + 64 * jj;
        if (c >= p && c <= K - 2) {
          const double y = rsqrt_refined(qn[jj]);                            // 1 / r_c
          const double a = (c == p) ? wp[jj] : q[jj] * rsqrt_refined(q[jj]);  // a_c
          s_v[2 * c] = wn[jj] * y;
          s_v[2 * c + 1] = a * y;
        }
      }
      shift_down<SPL>(aidx, lane, p, K - 1);
      shift_down<SPL>(lp, lane, p, K - 1);
      shift_down<SPL>(lv, lane, p, K - 1);
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj)
        if (lane + 64 * jj == K - 1) {
          aidx[jj] = -1;
          lp[jj] = lv[jj] = 0.0;
        }
      wave_sync();
      // rows p..K-2 of the new inverse factor, column j (old column j, or j + 1 past p) per lane;
      // new row c is written over old row c, already read
      double cur[SPL];
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) {
        const int j = lane + 64 * jj;
        const double v = s_W[j < p ? T(p) + j : 0];
        cur[jj] = j < p ? v : 0.0;
      }
      int wb = T(p);
      for (int c = p; c <= K - 2; ++c) {
        const double cs = s_v[2 * c], sn = s_v[2 * c + 1];
        const int rb = wb + c + 1;   // start of old row c + 1
#pragma unroll
        for (int jj = 0; jj < SPL; ++jj) {
          const int j = lane + 64 * jj;
          if (j <= c) {
            const double nx = s_W[rb + (j < p ? j : j + 1)];
            s_W[wb + j] = fma(cs, cur[jj], -sn * nx);
            cur[jj] = fma(sn, cur[jj], cs * nx);
          }
        }
        wb = rb;
      }
      k = K - 1;
    };

    auto set_bit = [&](int r, bool on) __attribute__((always_inline)) {
      if (lane == (r >> 6)) {
        const unsigned long long m = 1ull << (r & 63);
        actw = on ? (actw | m) : (actw & ~m);
      }
    };

    bool overflow = false;
    for (int e = 0; e < P.num_eq; ++e) {
      if (k >= KCAP) { overflow = true; break; }
      append(num_ineq + e);
      set_bit(num_ineq + e, true);
    }

    int niter = 0;
    double c[KQ];
#pragma unroll
    for (int q = 0; q < KQ; ++q) c[q] = cu[q];
    while (!overflow) {
      // most violated inequality, lowest index on ties (:270)
      double bv = INFINITY;
      int bi = INT_MAX;
#pragma unroll
      for (int q = 0; q < KQ; ++q) {
        const int r = lane + 64 * q;
        if (r < num_ineq && c[q] < bv) { bv = c[q]; bi = r; }
      }
      wave_argmin(bv, bi);
      bi = uniform32(bi);
      if (!uniform32(bv < -P.tol ? 1 : 0)) break;
      const bool was_active = (readlane64(actw, bi >> 6) >> (bi & 63)) & 1ull;
      if (was_active) {
        // the reference solves once more with the same set, gets the same (feasible)
        // multipliers, counts the iteration and stops (:272-346)
        const unsigned long long kk = (unsigned long long)k;
        ++st_inner; st_k += kk; st_k2 += kk * kk; st_k3 += kk * kk * kk;
        ++niter;
        break;
      }
      if (k >= KCAP) { overflow = true; break; }
      append(bi);
      set_bit(bi, true);

      for (;;) {   // inner loop (:274-331); l is the solution for the current set
        {
          const unsigned long long kk = (unsigned long long)k;
          ++st_inner; st_k += kk; st_k2 += kk * kk; st_k3 += kk * kk * kk;
        }
        // worst offender among negative inequality multipliers (:297-313)
        double sv = INFINITY;
        int si = INT_MAX;
#pragma unroll
        for (int jj = 0; jj < SPL; ++jj) {
          const int s = lane + 64 * jj;
          if (s < k && aidx[jj] < num_ineq && lv[jj] < 0.0) {
            const double sj = lp[jj] / (lp[jj] - lv[jj]);
            if (sj < sv || (sj == sv && aidx[jj] < si)) { sv = sj; si = aidx[jj]; }
          }
        }
        wave_argmin(sv, si);
        si = uniform32(si);
        if (si == INT_MAX) break;   // all feasible: x = y_u + B_act^T l (:316-321)
        int p = 0;
        {
          unsigned long long hit[SPL];
#pragma unroll
          for (int jj = 0; jj < SPL; ++jj) hit[jj] = __ballot(lane + 64 * jj < k && aidx[jj] == si);
#pragma unroll
          for (int jj = SPL - 1; jj >= 0; --jj)
            if (hit[jj]) p = 64 * jj + __ffsll((long long)hit[jj]) - 1;
        }
        set_bit(si, false);
        remove(p);
      }

      // lambda_prev = lambda (:334): the multipliers of the active set, zero elsewhere
#pragma unroll
      for (int jj = 0; jj < SPL; ++jj) lp[jj] = (lane + 64 * jj < k) ? lv[jj] : 0.0;
      ++niter;
      if (niter > P.max_niter) break;

      // c = B x - t = c_u + G[:, act] l (:349-351)
      st_kc += (unsigned long long)k;
      stage_pairs();
#pragma unroll
      for (int q = 0; q < KQ; ++q) c[q] = cu[q];
#pragma unroll 4
      for (int s = 0; s < k; ++s) {
        const int as = (int)s_v[2 * s + 1];
        const double ls = s_v[2 * s];
        const double* g = P.G + (size_t)as * ldg + lane;
#pragma unroll
        for (int q = 0; q < KQ; ++q)
          c[q] = fma(g[64 * q], ls, c[q]);
      }
    }

    if (overflow && Q.out_list) {
      // solved again, from the start, by the next pass
      if (lane == 0) Q.out_list[atomicAdd(Q.out_count, 1ull)] = (uint32_t)vox;
      continue;
    }

    // ---- x = y_u + B_act^T l (:319), then L_H^T x = y solved in place (:355) ----
    stage_pairs();
    double y[2] = {yu[0], yu[1]};
    for (int s = 0; s < k; ++s) {
      const int as = (int)s_v[2 * s + 1];
      const double ls = s_v[2 * s];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        if (h == 0 || n > 64) {
          const double bt = P.Bt[(size_t)(j < n ? j : 0) * ldg + as];
          if (j < n) y[h] = fma(bt, ls, y[h]);
        }
      }
    }
    // back substitution by columns: step i finalises x_i and removes it from the rows above
    for (int i = n - 1; i >= 0; --i) {
      const double xi = slot_get<2>(y, i) / P.L[(size_t)i * n + i];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        const double lij = P.L[(size_t)i * n + (j < i ? j : 0)];
        if (j == i) y[h] = xi;
        else if (j < i) y[h] = fma(-lij, xi, y[h]);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = lane + 64 * h;
      if (j < n) {
        if constexpr (IO64) L.x64[vox * n + j] = y[h];
        else L.x[vox * n + j] = (float)y[h];
      }
    }
    if (lane == 0) {
      if (L.iters) L.iters[vox] = overflow ? (uint16_t)0xFFFE : (uint16_t)niter;
      if (L.nactive) L.nactive[vox] = (uint16_t)k;
    }
    ++st_vox;
    st_outer += (unsigned long long)niter;
    st_ovf += overflow ? 1ull : 0ull;
  }

  if (L.stats && lane == 0) {
    const unsigned long long v[8] = {st_vox, st_outer, st_inner, st_k, st_k2, st_k3, st_kc, st_ovf};
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (v[i]) atomicAdd(L.stats + i, v[i]);
  }
}

}  // namespace

int icls_default_capacity(int n, int num_eq) {
  for (int i = 0; i < kIclsNumCapacities; ++i)
    if (kIclsCapacities[i] > n + num_eq) return kIclsCapacities[i];
  return kIclsCapacities[kIclsNumCapacities - 1];
}

int icls_constraint_words(int ncon) {
  const int w = (ncon + 63) / 64;
  return w <= 4 ? 4 : w <= 5 ? 5 : w <= 8 ? 8 : 16;
}

bool icls_valid_capacity(int kcap) {
  for (int i = 0; i < kIclsNumCapacities; ++i)
    if (kIclsCapacities[i] == kcap) return true;
  return false;
}

namespace {

// The passes of a solve: the chosen capacity, then 64 and 128 slots for the voxels that outgrow
// the previous one.  Returns their number.
int plan_passes(int first, int (&caps)[kIclsMaxPasses]) {
  int npass = 0;
  caps[npass++] = first;
  if (first < 64) caps[npass++] = 64;
  if (first < 128) caps[npass++] = 128;
  return npass;
}

// Queues and counters of pass i of npass: it reads the queue the previous pass wrote.
IclsPass pass_args(const IclsScratch& S, int i, int npass) {
  IclsPass Q{};
  Q.counter = S.counters + i;
  if (i > 0) {
    Q.in_list = S.queue[(i - 1) & 1];
    Q.in_count = S.counters + kIclsMaxPasses + (i - 1);
  }
  if (i + 1 < npass) {
    Q.out_list = S.queue[i & 1];
    Q.out_count = S.counters + kIclsMaxPasses + i;
  }
  return Q;
}

// ---- HIP launches ----

__global__ void reset_counters_kernel(unsigned long long* c, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

template <int KCAP, int KQ>
hipError_t launch_pass(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsPass& Q, size_t max_grid,
                       hipStream_t stream) {
  auto kern = L.b64 ? icls_active_set_kernel<KCAP, KQ, true> : icls_active_set_kernel<KCAP, KQ, false>;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  size_t grid = (size_t)ncu * (size_t)per_cu;
  if (grid > max_grid) grid = max_grid;
  if (grid < 1) grid = 1;
  kern<<<dim3((unsigned)grid), dim3(64), 0, stream>>>(P, L, Q);
  return hipGetLastError();
}

template <int KCAP>
hipError_t launch_kq(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsPass& Q, size_t max_grid,
                     hipStream_t stream) {
  switch (P.ncon_pad / 64) {
    case 4: return launch_pass<KCAP, 4>(P, L, Q, max_grid, stream);
    case 5: return launch_pass<KCAP, 5>(P, L, Q, max_grid, stream);
    case 8: return launch_pass<KCAP, 8>(P, L, Q, max_grid, stream);
    case 16: return launch_pass<KCAP, 16>(P, L, Q, max_grid, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_capacity(int kcap, const IclsDeviceProblem& P, const IclsLaunch& L, const IclsPass& Q,
                           size_t max_grid, hipStream_t stream) {
  switch (kcap) {
    case 16: return launch_kq<16>(P, L, Q, max_grid, stream);
    case 32: return launch_kq<32>(P, L, Q, max_grid, stream);
    case 48: return launch_kq<48>(P, L, Q, max_grid, stream);
    case 64: return launch_kq<64>(P, L, Q, max_grid, stream);
    case 96: return launch_kq<96>(P, L, Q, max_grid, stream);
    case 128: return launch_kq<128>(P, L, Q, max_grid, stream);
    default: return hipErrorInvalidValue;
  }
}

// Four voxels per workgroup: their solutions in LDS, a thread per (voxel, measurement) output, so
// the reads of Ht and the writes of out are consecutive across threads.
constexpr int kPredictVox = 4;

__global__ __launch_bounds__(256) void icls_predict_kernel(const double* __restrict__ Ht, int n, int nm,
                                                           const float* __restrict__ x, int ld_x, size_t nvox,
                                                           float* __restrict__ out, int ld_out) {
  __shared__ double xs[kPredictVox][kIclsMaxParams];
  const size_t v0 = (size_t)blockIdx.x * kPredictVox;
  for (int i = threadIdx.x; i < kPredictVox * n; i += 256) {
    const int vl = i / n, j = i - vl * n;
    const size_t v = v0 + vl;
    xs[vl][j] = v < nvox ? (double)x[v * (size_t)ld_x + j] : 0.0;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kPredictVox * nm; q += 256) {
    const int vl = q / nm, r = q - vl * nm;
    const size_t v = v0 + vl;
    if (v >= nvox) continue;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = fma(Ht[(size_t)j * nm + r], xs[vl][j], acc);
    out[v * (size_t)ld_out + r] = (float)acc;
  }
}

}  // namespace

hipError_t icls_predict(const double* Ht, int n, int nmeas, const float* x, int ld_x, size_t nvox, float* out,
                        int ld_out, hipStream_t stream) {
  if (n < 1 || n > kIclsMaxParams || nmeas < 1 || ld_x < n || ld_out < nmeas) return hipErrorInvalidValue;
  if (nvox == 0) return hipSuccess;
  const size_t blocks = (nvox + kPredictVox - 1) / kPredictVox;
  icls_predict_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(Ht, n, nmeas, x, ld_x, nvox, out, ld_out);
  return hipGetLastError();
}

hipError_t icls_launch(const IclsDeviceProblem& P, const IclsLaunch& L, const IclsScratch& S,
                       int first_capacity, hipStream_t stream) {
  if (!icls_valid_capacity(first_capacity) || P.ncon_pad != 64 * icls_constraint_words(P.ncon))
    return hipErrorInvalidValue;
  int caps[kIclsMaxPasses];
  const int npass = plan_passes(first_capacity, caps);
  if (npass > 1 && S.capacity < L.nvox) return hipErrorInvalidValue;
  reset_counters_kernel<<<1, 64, 0, stream>>>(S.counters, 2 * kIclsMaxPasses);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int i = 0; i < npass; ++i) {
    // later passes see few voxels: a small grid that exits at once when its queue is empty
    const size_t max_grid = i == 0 ? (L.nvox ? L.nvox : 1) : 256;
    if ((e = launch_capacity(caps[i], P, L, pass_args(S, i, npass), max_grid, stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mrcsd
