// Gauss-Jordan elimination with one matrix row per thread, as the CSD kernel solves its normal equations
// (gfx950, wavefront size 64).  Thread i owns row i of a symmetric positive definite N x N system
// [A | rhs]; after the sweep A is diagonal and x_i = rhs_i / di.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "wave_utils.h"

namespace mrcsd {
namespace gj {

// doubles per pivot-column buffer: NT column entries, then (workgroups wider than one wavefront)
// the NT right-hand sides, so that the pivot's right-hand side is read from LDS.
template <int NT>
constexpr int col_stride() { return NT == 64 ? NT + 2 : 2 * NT; }

// The lane index as a value the compiler cannot treat as loop-invariant: per-column masks and
// addresses derived from it are recomputed where they are used instead of being hoisted out of
// the persistent voxel loop (where 45 of them would stay live in registers).
__device__ __forceinline__ int opaque_lane() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// A pointer into LDS whose value the compiler cannot see, held in one VGPR (the low 32 bits of a
// generic pointer into LDS are its LDS address; converting this way avoids the null check of an
// address-space cast).  Loads through it use immediate offsets, so a chunk of the pivot column
// costs one address move, and loads through different such pointers are not merged into one
// batch issued up front.
typedef __attribute__((address_space(3))) double lds_f64;
__device__ __forceinline__ const lds_f64* opaque_lds(const double* p) {
  unsigned a = (unsigned)(size_t)p;
  asm volatile("" : "+v"(a));
  return (const lds_f64*)(size_t)a;
}

constexpr int kPipe = 4;   // elements per software-pipelined load chunk

// The hot loops below are software pipelined by hand: the loads of chunk k+1 are issued before the
// FMAs of chunk k (double buffering), and sched_barrier keeps the compiler from moving them back
// next to their uses.  Left alone, the scheduler places every load directly before its FMA and
// each FMA then waits for a full LDS or L2 round trip.

// One Gauss-Jordan step on [A | rhs], thread i owning row i: eliminates column KK from every
// row but KK.  The pivot row past the diagonal equals column KK (symmetric trailing block), read
// from `cur`; the next step's column is updated first and published to `nxt` (look-ahead) so its
// LDS round trip overlaps the remaining FMAs.  The first chunk of the pivot column is in flight
// during the reciprocal.  A step has no branches (the reciprocal runs on every lane and is masked;
// every lane publishes): a conditional block per step lets the compiler sink each step's FMAs
// into the next one, which keeps all the loaded pivot columns live and spills.  The pivot row's
// diagonal is taken into `di` at its own step, so columns already eliminated are dead afterwards
// and the live part of the row shrinks as the sweep proceeds.
template <int NSH, int NT, int KK>
__device__ __forceinline__ void gj_step(double (&A)[NSH], double& rhs, double& di, double* col0, double* col1,
                                        int tid) {
  double* const cur = (KK & 1) ? col1 : col0;
  double* const nxt = (KK & 1) ? col0 : col1;
  constexpr int J0 = KK + 1;
  double d, bk;
  if constexpr (NT == 64) {
    d = lane_bcast(A[KK], KK);
    bk = lane_bcast(rhs, KK);
  } else {
    d = cur[KK];    // published by lane KK in the previous step (or before step 0)
    bk = cur[NT + KK];
  }
  double buf[2][kPipe];
  if constexpr (J0 < NSH) {
    const lds_f64* p = opaque_lds(cur + J0);
#pragma unroll
    for (int i = 0; i < kPipe; ++i)
      if (J0 + i < NSH) buf[0][i] = p[i];
  }
  __builtin_amdgcn_sched_barrier(0);   // the first chunk's loads are issued before the reciprocal
  const int t = opaque_lane();   // per-step lane masks are recomputed, not hoisted (one SGPR pair each)
  // A[KK] / d through the pivot's reciprocal: v_rcp_f64 refined by two Newton steps (about one ulp;
  // a correctly rounded division is a longer sequence, and every lane executes it)
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  const double q = A[KK] * r;
  const double c = (t != KK) ? q : 0.0;
  di = (t == KK) ? A[KK] : di;
  const double rn = fma(-c, bk, rhs);
  if constexpr (J0 < NSH) {   // the last step has no trailing columns and no successor
#pragma unroll
    for (int k = 0; J0 + k * kPipe < NSH; ++k) {
      const int j = J0 + k * kPipe;
      if (j + kPipe < NSH) {
        const lds_f64* p = opaque_lds(cur + j + kPipe);
#pragma unroll
        for (int i = 0; i < kPipe; ++i)
          if (j + kPipe + i < NSH) buf[(k + 1) & 1][i] = p[i];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < kPipe; ++i)
        if (j + i < NSH) A[j + i] = fma(-c, buf[k & 1][i], A[j + i]);
      if (k == 0) {
        nxt[tid] = A[J0];
        if constexpr (NT != 64) nxt[NT + tid] = rn;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  rhs = rn;
  wave_sync();
}

template <int NSH, int NT, int... KK>
__device__ __forceinline__ void gj_sweep(double (&A)[NSH], double& rhs, double& di, double* col0, double* col1,
                                         int tid, std::integer_sequence<int, KK...>) {
  (gj_step<NSH, NT, KK>(A, rhs, di, col0, col1, tid), ...);
}

}  // namespace gj
}  // namespace mrcsd
