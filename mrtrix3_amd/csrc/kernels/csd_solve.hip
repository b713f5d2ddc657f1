// Batched constrained spherical deconvolution on gfx950 (MI355X).
//
// One workgroup (NT threads = NSH rounded up to whole wavefronts) solves one voxel at a time and
// pulls the next voxel from a global work counter when it finishes, so voxels that converge after
// different numbers of iterations retire and refill without a relaunch (persistent kernel).
//
// Per voxel this is CSD::set + the dwi2fod iterate loop (src/dwi/sdeconv/csd.h:236-269,
// cmd/dwi2fod.cpp:128-173), all in FP64:
//   set      : Mt_b = M^T b and F = [rconv b; 0]                     (thread i owns coefficient i)
//   iterate  : a = HR_trans F over the constraint directions, neg = {r : a_r < threshold}
//              (ballots build a bitmask); unchanged neg set => converged.
//              W = Mt_M + sum_{r in neg} h_r h_r^T (thread i owns row i while solving; the
//              packed lower triangle lives in LDS between iterations) is updated only by the
//              directions that entered or left neg since the last solve; then W F = Mt_b is solved by
//              Gauss-Jordan elimination with diagonal pivots: thread i eliminates column k from
//              its own row for every k != i, reading the pivot row k from LDS.  Because the
//              trailing block of a symmetric matrix stays symmetric under elimination, pivot row
//              k equals column k there, which every thread writes in one LDS store.  After the
//              sweep F_i = b_i / a_ii, so no triangular factor is stored and no backward pass runs.
// The reference solves with Cholesky LL^T (Eigen::LLT).  Gauss-Jordan without pivoting is
// backward stable for symmetric positive definite W; the solutions agree to rounding, not bit
// for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "csd_solve.h"

namespace mrcsd {
namespace {

constexpr int kMaxDirWords = 8;   // up to 512 constraint directions
constexpr int kMaxDW = 256;       // up to 256 DW volumes per voxel
constexpr int kMaxQ = 8;          // constraint directions per thread (n_hr_pad / NT)
constexpr int kChunk = 8;         // normal-matrix update rows staged per LDS hand-off
constexpr int kColStride = 66;    // doubles per pivot-column buffer (16-byte aligned)

// A value the whole wavefront agrees on, moved to scalar registers so that loops over it are
// scalar loops.
__device__ __forceinline__ unsigned long long uniform64(unsigned long long x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(x & 0xffffffffull));
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Broadcast lane `src`'s double to the whole wavefront (src must be wave-uniform).
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffull), src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The workgroup is a single wavefront, so LDS hand-offs only need the compiler to keep program
// order and wait for the LDS queue; __syncthreads() provides both (the barrier itself is elided
// for one-wave workgroups).
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// The lane index as a value the compiler cannot treat as loop-invariant: per-column masks and
// addresses derived from it are recomputed where they are used instead of being hoisted out of
// the persistent voxel loop (where 45 of them would stay live in registers).
__device__ __forceinline__ int opaque_lane() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Row `i` of a symmetric matrix whose lower triangle is packed row by row (element (r,c), c <= r,
// at r(r+1)/2 + c): the lower part comes from row i, the upper part from column i of later rows.
template <int NSH>
__device__ __forceinline__ void load_sym_row(const double* W, double (&A)[NSH]) {
  const int t = opaque_lane();
  const int tt = t < NSH ? t : NSH - 1;
  const double* lo = W + tt * (tt + 1) / 2;
  const double* up = W + tt;
#pragma unroll
  for (int j = 0; j < NSH; ++j) {
    const double a = lo[j <= tt ? j : 0];
    const double b = up[j * (j + 1) / 2];
    A[j] = (j <= tt) ? a : b;
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);   // bounds the loads in flight
  }
}

// Stores the lower part of row i (columns 0..i) into the packed triangle.
template <int NSH>
__device__ __forceinline__ void store_lower_row(double* W, const double (&A)[NSH]) {
  const int t = opaque_lane();
  if (t < NSH) {
    double* lo = W + t * (t + 1) / 2;
#pragma unroll
    for (int j = 0; j < NSH; ++j)
      if (j <= t) lo[j] = A[j];
  }
}

// One Gauss-Jordan step on [A | rhs], thread i owning row i: eliminates column KK from every
// row but KK.  The pivot row past the diagonal equals column KK (symmetric trailing block), read
// from `cur`; the next step's column is updated first and published to `nxt` (look-ahead) so its
// LDS round trip overlaps the remaining FMAs.
template <int NSH, int KK>
__device__ __forceinline__ void gj_step(double (&A)[NSH], double& rhs, double* col0, double* col1, int tid) {
  double* const cur = (KK & 1) ? col1 : col0;
  double* const nxt = (KK & 1) ? col0 : col1;
  const double d = lane_bcast(A[KK], KK);
  const double bk = lane_bcast(rhs, KK);
  const double c = (tid != KK) ? A[KK] / d : 0.0;
  if constexpr (KK + 1 < NSH) {
    A[KK + 1] = fma(-c, cur[KK + 1], A[KK + 1]);
    nxt[tid] = A[KK + 1];
  }
#pragma unroll
  for (int j = KK + 2; j < NSH; ++j) {
    A[j] = fma(-c, cur[j], A[j]);
    if (((j - KK) & 7) == 7) __builtin_amdgcn_sched_barrier(0);   // bounds the loads in flight
  }
  rhs = fma(-c, bk, rhs);
  wave_sync();
}

template <int NSH, int... KK>
__device__ __forceinline__ void gj_sweep(double (&A)[NSH], double& rhs, double* col0, double* col1, int tid,
                                         std::integer_sequence<int, KK...>) {
  (gj_step<NSH, KK>(A, rhs, col0, col1, tid), ...);
}

// amdgpu_waves_per_eu(3): caps the kernel at 168 VGPRs so three wavefronts share each SIMD
// (the LDS footprint, ~12.8 KB per wavefront at NSH = 45, allows twelve per CU).
template <int NSH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void csd_persistent_kernel(CsdDeviceProblem P, const float* __restrict__ dw,
                                                            int ld_dw, const uint8_t* __restrict__ valid,
                                                            size_t nvox, float* __restrict__ fod,
                                                            uint16_t* __restrict__ iters_out,
                                                            uint16_t* __restrict__ nneg_out,
                                                            unsigned long long* __restrict__ stats,
                                                            const float* __restrict__ warm_F0,
                                                            unsigned long long* __restrict__ negmask,
                                                            unsigned long long* __restrict__ counter) {
  static_assert(NSH <= 64, "one wavefront owns one row per lane");
  constexpr int NT = 64;
  constexpr int NTRI = NSH * (NSH + 1) / 2;
  constexpr int COLS = 2 * kColStride;                       // two pivot-column buffers
  constexpr int HBUF = kChunk * NSH > COLS ? kChunk * NSH : COLS;
  __shared__ float s_b[kMaxDW];
  __shared__ double s_F[NT];
  __shared__ __attribute__((aligned(16))) double s_h[HBUF];  // update rows, then pivot columns
  __shared__ double s_W[NTRI];        // packed lower triangle of the normal matrix
  __shared__ unsigned long long s_words[kMaxDirWords];

  const int tid = threadIdx.x;
  const int N = P.n_dw;
  const int R = P.n_hr;
  const int Rpad = P.n_hr_pad;
  const int nq = Rpad / NT;   // host guarantees Rpad % 64 == 0 and Rpad / 64 <= kMaxDirWords
  const bool row_ok = tid < NSH;

  unsigned long long st_vox = 0, st_solves = 0, st_k = 0, st_conv = 0;

  for (;;) {
    unsigned long long vox = 0;
    if (tid == 0) vox = atomicAdd(counter, 1ull);
    vox = uniform64(vox);
    if (vox >= nvox) break;

    // ---- gather the selected DW signals (dwi2fod.cpp:156-173) ----
    bool bad = valid && !valid[vox];
    const float* src = dw + vox * (size_t)ld_dw;
    for (int r = tid; r < N; r += NT) {
      const float v = src[P.dwis[r]];
      if (!isfinite(v)) bad = true;
      s_b[r] = v < 0.0f ? 0.0f : v;
    }
    if (__ballot(bad)) {
      if (row_ok) fod[vox * NSH + tid] = 0.0f;
      if (tid == 0) {
        if (iters_out) iters_out[vox] = 0xFFFF;
        if (nneg_out) nneg_out[vox] = 0;
      }
      if (negmask && tid < nq) negmask[vox * nq + tid] = 0ull;
      continue;
    }
    wave_sync();

    // ---- CSD::set: Mt_b and the initial F (csd.h:236-243) ----
    double mtb = 0.0, f = 0.0;
    if (row_ok) {
      for (int r = 0; r < N; ++r) mtb = fma(P.M[r * NSH + tid], (double)s_b[r], mtb);
      if (warm_F0) {
        f = (double)warm_F0[vox * NSH + tid];
      } else if (tid < P.n_init) {
        for (int r = 0; r < N; ++r) f = fma(P.rconvT[r * P.n_init + tid], (double)s_b[r], f);
      }
    }
    {
      double A0[NSH];
      const double* mrow = P.MtM + (row_ok ? tid : 0) * NSH;
#pragma unroll
      for (int j = 0; j < NSH; ++j) A0[j] = mrow[j];
      store_lower_row<NSH>(s_W, A0);
    }

    unsigned long long oldw[kMaxDirWords];
#pragma unroll
    for (int w = 0; w < kMaxDirWords; ++w) oldw[w] = 0ull;
    bool have_old = false;   // old_neg = [-1] (csd.h:240)
    int n = 0;
    int last_k = 0;

    for (n = 0; n < P.niter; ++n) {
      // ---- amplitudes and the negative set (csd.h:246-250) ----
      s_F[tid] = f;
      wave_sync();
      {
        double a[kMaxQ];
#pragma unroll
        for (int q = 0; q < kMaxQ; ++q) a[q] = 0.0;
        const double* hp = P.HRt + tid;
#pragma unroll 3
        for (int j = 0; j < NSH; ++j) {
          const double fj = s_F[j];
#pragma unroll
          for (int q = 0; q < kMaxQ; ++q)
            if (q < nq) a[q] = fma(hp[q * NT], fj, a[q]);
          hp += Rpad;
          __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int q = 0; q < kMaxQ; ++q) {
          if (q < nq) {
            const int r = tid + q * NT;
            const unsigned long long ball = __ballot(r < R && a[q] < P.thr);
            if (tid == 0) s_words[q] = ball;
          }
        }
      }
      wave_sync();
      bool same = have_old;
      int k = 0;
      unsigned long long neww[kMaxDirWords];
#pragma unroll
      for (int w = 0; w < kMaxDirWords; ++w) {
        neww[w] = w < nq ? uniform64(s_words[w]) : 0ull;
        k += __popcll(neww[w]);
        same = same && (neww[w] == oldw[w]);
      }
      if (same) break;   // converged (csd.h:252-253)

      // ---- W = Mt_M + H_neg^T H_neg (csd.h:255-262), updated by the change of the neg set ----
      // Row i of the symmetric W is read from the packed lower triangle (element (i,j) sits at
      // max(i,j)(max(i,j)+1)/2 + min(i,j)).  Directions that entered (+) or left (-) neg are
      // applied in chunks of kChunk rows staged in LDS.
      double A[NSH];
      load_sym_row<NSH>(s_W, A);
      {
        int w = 0;
        unsigned long long delta = have_old ? (neww[0] ^ oldw[0]) : neww[0];
        for (;;) {
          int cnt = 0;
          int rows_c[kChunk];
          double sgn_c[kChunk];
#pragma unroll
          for (int c = 0; c < kChunk; ++c) {
            while (delta == 0ull && w + 1 < nq) {
              ++w;
              unsigned long long nw = 0ull, ow = 0ull;
#pragma unroll
              for (int q = 0; q < kMaxDirWords; ++q) {
                nw = (q == w) ? neww[q] : nw;
                ow = (q == w) ? oldw[q] : ow;
              }
              delta = have_old ? (nw ^ ow) : nw;
            }
            rows_c[c] = 0;
            sgn_c[c] = 0.0;
            if (delta) {
              const int bit = __ffsll((long long)delta) - 1;
              delta &= delta - 1;
              unsigned long long nw = 0ull;
#pragma unroll
              for (int q = 0; q < kMaxDirWords; ++q) nw = (q == w) ? neww[q] : nw;
              rows_c[c] = w * 64 + bit;
              sgn_c[c] = ((nw >> bit) & 1ull) ? 1.0 : -1.0;
              ++cnt;
            }
          }
          if (cnt == 0) break;
          double hown[kChunk];
#pragma unroll
          for (int c = 0; c < kChunk; ++c)
            hown[c] = (c < cnt) ? P.HR[(size_t)rows_c[c] * NSH + (row_ok ? tid : 0)] : 0.0;
          wave_sync();   // earlier readers of s_h are done
          if (row_ok) {
#pragma unroll
            for (int c = 0; c < kChunk; ++c) s_h[c * NSH + tid] = hown[c];
          }
          wave_sync();
#pragma unroll
          for (int c = 0; c < kChunk; ++c) {
            if (c < cnt) {
              const double hi = sgn_c[c] * hown[c];
#pragma unroll
              for (int j = 0; j < NSH; ++j) {
                A[j] = fma(hi, s_h[c * NSH + j], A[j]);
                if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
              }
            }
          }
          if (cnt < kChunk) break;
        }
      }
      store_lower_row<NSH>(s_W, A);
#pragma unroll
      for (int w = 0; w < kMaxDirWords; ++w) oldw[w] = neww[w];
      have_old = true;
      last_k = k;
      ++st_solves;
      st_k += (unsigned long long)k;

      // ---- Gauss-Jordan sweep on [W | Mt_b] (thread i owns row i) ----
      // Step kk eliminates column kk from every other row.  Pivot row kk equals column kk past
      // the diagonal, so each lane publishes its A[kk] as column kk.  The next column is
      // updated first and published before the rest of the row (look-ahead), so its LDS
      // round trip overlaps the remaining FMAs.
      double* const col0 = s_h;
      double* const col1 = s_h + kColStride;
      double rhs = mtb;
      wave_sync();   // s_h (update rows) is free again
      col0[tid] = A[0];
      wave_sync();
      gj_sweep<NSH>(A, rhs, col0, col1, tid, std::make_integer_sequence<int, NSH>{});
      double di = 1.0;
#pragma unroll
      for (int j = 0; j < NSH; ++j) di = (tid == j) ? A[j] : di;
      f = row_ok ? rhs / di : 0.0;
    }

    if (row_ok) fod[vox * NSH + tid] = (float)f;
    if (tid == 0) {
      if (iters_out) iters_out[vox] = (uint16_t)n;
      if (nneg_out) nneg_out[vox] = (uint16_t)last_k;
    }
    if (negmask) {
      unsigned long long mine = 0ull;
#pragma unroll
      for (int w = 0; w < kMaxDirWords; ++w) mine = (tid == w) ? oldw[w] : mine;
      if (tid < nq) negmask[vox * nq + tid] = mine;
    }
    ++st_vox;
    if (n < P.niter) ++st_conv;
    wave_sync();
  }

  if (stats && tid == 0) {
    atomicAdd(stats + 0, st_vox);
    atomicAdd(stats + 1, st_solves);
    atomicAdd(stats + 2, st_k);
    atomicAdd(stats + 3, st_conv);
  }
}

__global__ void reset_counter_kernel(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

template <int NSH>
hipError_t launch_impl(const CsdDeviceProblem& P, const CsdLaunch& L, hipStream_t stream) {
  constexpr int NT = 64;
  auto kern = csd_persistent_kernel<NSH>;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  int dev = 0, ncu = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  size_t grid = (size_t)ncu * (size_t)per_cu;
  if (L.max_groups > 0 && grid > (size_t)L.max_groups) grid = (size_t)L.max_groups;
  if (grid > L.nvox) grid = L.nvox ? L.nvox : 1;
  reset_counter_kernel<<<1, 64, 0, stream>>>(L.counter);
  kern<<<dim3((unsigned)grid), dim3(NT), 0, stream>>>(P, L.dw, L.ld_dw, L.valid, L.nvox, L.fod, L.iters,
                                                     L.nneg, L.stats, L.warm_F0, L.negmask, L.counter);
  return hipGetLastError();
}

}  // namespace

hipError_t csd_launch(const CsdDeviceProblem& P, const CsdLaunch& L, hipStream_t stream) {
  switch (P.n_sh) {
    case 6:  return launch_impl<6>(P, L, stream);
    case 15: return launch_impl<15>(P, L, stream);
    case 28: return launch_impl<28>(P, L, stream);
    case 45: return launch_impl<45>(P, L, stream);
    default: return hipErrorInvalidValue;
  }
}

bool csd_supported_nsh(int n_sh) {
  return n_sh == 6 || n_sh == 15 || n_sh == 28 || n_sh == 45;
}

int csd_threads_for(int) { return 64; }

}  // namespace mrcsd
