// Batched constrained spherical deconvolution on gfx950 (MI355X).
//
// One workgroup (NT = 64 threads for up to 64 SH coefficients, 128 for lmax 10-12) solves one
// voxel at a time and
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
#include "gauss_jordan.h"
#include "wave_utils.h"

namespace mrcsd {
namespace {

using gj::col_stride;
using gj::gj_sweep;
using gj::kPipe;
using gj::opaque_lane;

constexpr int kMaxDirWords = 8;   // up to 512 constraint directions
constexpr int kMaxDW = 256;       // up to 256 DW volumes per voxel
// Row `i` of a symmetric matrix whose lower triangle is packed row by row (element (r,c), c <= r,
// at r(r+1)/2 + c): the lower part comes from row i, the upper part from column i of later rows.
template <int NSH>
__device__ __forceinline__ void load_sym_row(const double* W, double (&A)[NSH]) {
  const int t = opaque_lane();
  const int tt = t < NSH ? t : NSH - 1;
  const double* const lo = W + tt * (tt + 1) / 2;
  const double* const up = W + tt;
  double buf[2][kPipe];
#pragma unroll
  for (int i = 0; i < kPipe; ++i)
    if (i < NSH) buf[0][i] = *((i <= tt) ? lo + i : up + i * (i + 1) / 2);
#pragma unroll
  for (int k = 0; k * kPipe < NSH; ++k) {
    const int jn = (k + 1) * kPipe;
    if (jn < NSH) {
#pragma unroll
      for (int i = 0; i < kPipe; ++i) {
        const int j = jn + i;
        if (j < NSH) buf[(k + 1) & 1][i] = *((j <= tt) ? lo + j : up + j * (j + 1) / 2);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < kPipe; ++i)
      if (k * kPipe + i < NSH) A[k * kPipe + i] = buf[k & 1][i];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Stores the lower part of row i (columns 0..i) into the packed triangle.  Branch free: the other
// columns, and threads without a row, write to their own slot of `junk` instead.
template <int NSH>
__device__ __forceinline__ void store_lower_row(double* W, double* junk, const double (&A)[NSH]) {
  const int t = opaque_lane();
  const int tt = t < NSH ? t : -1;
  double* const lo = W + (t < NSH ? t * (t + 1) / 2 : 0);
  double* const jk = junk + t;
#pragma unroll
  for (int j = 0; j < NSH; ++j) *((j <= tt) ? lo + j : jk) = A[j];
}

typedef double v4d __attribute__((ext_vector_type(4)));

// The normal matrix as 16x16 tiles (n rounded up to whole tiles).  Lower tiles are numbered
// row-major over the lower triangle, g = I(I+1)/2 + J with J <= I; a workgroup of NWAVE
// wavefronts gives tile g to wavefront g % NWAVE.
constexpr int tile_row(int g) {
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= g) ++I;
  return I;
}
constexpr int tile_col(int g) { return g - tile_row(g) * (tile_row(g) + 1) / 2; }

template <int NSH, int NWAVE, int WAVE>
struct WaveTiles {
  static constexpr int NTL = (NSH + 15) / 16;
  static constexpr int NLOW = NTL * (NTL + 1) / 2;
  static constexpr int MYT = (NLOW - WAVE + NWAVE - 1) / NWAVE;
  static constexpr int g(int t) { return WAVE + t * NWAVE; }
};

// Element (row, col) of tile (I, J) held by lane l in register r of an FP64 16x16x4 MFMA
// accumulator: col = 16J + l % 16, row = 16I + l / 16 + 4r.  Its index in the packed lower
// triangle (element (a, b), b <= a, at a(a+1)/2 + b) mirrors it into the lower part on diagonal
// tiles; `ok` is false for padding and, when storing, for the upper half of diagonal tiles.
template <int NSH>
__device__ __forceinline__ int packed_index(int I, int J, int l, int r, bool store, bool& ok) {
  const int row = 16 * I + (l >> 4) + 4 * r;
  const int col = 16 * J + (l & 15);
  const int a = row > col ? row : col;
  const int b = row > col ? col : row;
  ok = row < NSH && col < NSH && (!store || I != J || row >= col);
  return a * (a + 1) / 2 + b;
}

// W += sum over the chunk's directions d of s_d h_d h_d^T on this wavefront's lower tiles, where
// h_d is row d of HR (the constraint direction's SH values) and s_d = +1 (entered the negative
// set) or -1 (left it).  One v_mfma_f64_16x16x4f64 per tile per four directions: the A operand
// (lane i + 16c holds s_c h_c[16I + i]) and the B operand (lane j + 16c holds h_c[16J + j]) come
// from the same loads.  The next chunk's loads are issued before the current chunk's MFMAs.
// take4(r0..r3, s0..s3) yields the next up-to-four directions (HR rows and signs, unused slots row
// 0 and sign 0; returns the count) and is evaluated identically by every wavefront of the
// workgroup.
template <int NSH, int NWAVE, int WAVE, class Take>
__device__ __forceinline__ void mfma_update(double* W, double* junk, const double* __restrict__ HR, Take& take4) {
  using G = WaveTiles<NSH, NWAVE, WAVE>;
  constexpr int NTL = G::NTL;
  const int l = opaque_lane() & 63;
  v4d acc[G::MYT];
#pragma unroll
  for (int t = 0; t < G::MYT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bool ok;
      const int idx = packed_index<NSH>(tile_row(G::g(t)), tile_col(G::g(t)), l, r, false, ok);
      const double x = W[ok ? idx : 0];
      acc[t][r] = ok ? x : 0.0;
    }
  }
  auto fetch = [&](int r0, int r1, int r2, int r3, double s0, double s1, double s2, double s3, double (&h)[NTL],
                   double& s) __attribute__((always_inline)) {
    const int c = l >> 4;
    const int row = c == 0 ? r0 : (c == 1 ? r1 : (c == 2 ? r2 : r3));
    s = c == 0 ? s0 : (c == 1 ? s1 : (c == 2 ? s2 : s3));
    const double* src = HR + (size_t)row * NSH + (l & 15);
#pragma unroll
    for (int T = 0; T < NTL; ++T) {
      const bool in = 16 * T + (l & 15) < NSH;
      const double x = src[in ? 16 * T : 0];
      h[T] = in ? x : 0.0;
    }
  };
  double h[NTL], s = 0.0;
  int cnt;
  {
    int r0, r1, r2, r3;
    double s0, s1, s2, s3;
    cnt = take4(r0, r1, r2, r3, s0, s1, s2, s3);
    fetch(r0, r1, r2, r3, s0, s1, s2, s3, h, s);
  }
  while (cnt > 0) {
    double h2[NTL], s2 = 0.0;
    int cnt2 = 0;
    if (cnt == 4) {
      int r0, r1, r2, r3;
      double q0, q1, q2, q3;
      cnt2 = take4(r0, r1, r2, r3, q0, q1, q2, q3);
      fetch(r0, r1, r2, r3, q0, q1, q2, q3, h2, s2);
    } else {
#pragma unroll
      for (int T = 0; T < NTL; ++T) h2[T] = 0.0;
    }
    __builtin_amdgcn_sched_barrier(0);
    double a[NTL];
#pragma unroll
    for (int T = 0; T < NTL; ++T) a[T] = s * h[T];
#pragma unroll
    for (int t = 0; t < G::MYT; ++t)
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tile_row(G::g(t))], h[tile_col(G::g(t))], acc[t], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int T = 0; T < NTL; ++T) h[T] = h2[T];
    s = s2;
    cnt = cnt2;
  }
#pragma unroll
  for (int t = 0; t < G::MYT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bool ok;
      const int idx = packed_index<NSH>(tile_row(G::g(t)), tile_col(G::g(t)), l, r, true, ok);
      *(ok ? W + idx : junk + l) = acc[t][r];
    }
  }
}

// amdgpu_waves_per_eu: caps the 64-thread kernel at 128 VGPRs so four wavefronts share each SIMD
// (the LDS footprint, ~9.9 KB per wavefront at NSH = 45, allows sixteen per CU); the kernel is
// latency bound, and the fourth wavefront hides more than the few spilled per-voxel values cost.
// The 128-thread kernel keeps 91-wide rows in registers and runs two wavefronts per SIMD.
// WARM: the launch carries a proposal (warm_F0 and/or warm_neg).  The plain instantiation has no
// warm pass, so nothing has to stay live across one; the warm instantiation forms rconv b only
// when a voxel falls back to the plain start.
template <int NSH, int NT, bool WARM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 64 ? 4 : 2))) void csd_persistent_kernel(CsdDeviceProblem P, const float* __restrict__ dw,
                                                            int ld_dw, const uint8_t* __restrict__ valid,
                                                            size_t nvox, float* __restrict__ fod,
                                                            uint16_t* __restrict__ iters_out,
                                                            uint16_t* __restrict__ nneg_out,
                                                            unsigned long long* __restrict__ stats,
                                                            const float* __restrict__ warm_F0,
                                                            const unsigned long long* __restrict__ warm_neg,
                                                            int warm_budget,
                                                            unsigned long long* __restrict__ negmask,
                                                            unsigned long long* __restrict__ counter) {
  static_assert(NSH <= NT && (NT == 64 || NT == 128), "one thread owns one row");
  constexpr int NWAVE = NT / 64;
  constexpr int NTRI = NSH * (NSH + 1) / 2;
  constexpr int kColStride = col_stride<NT>();
  constexpr int COLS = 2 * kColStride;                       // two pivot-column buffers
  constexpr int HBUF = COLS > kMaxDW / 2 ? COLS : kMaxDW / 2;
  __shared__ double s_F[NT];
  __shared__ __attribute__((aligned(16))) double s_h[HBUF];  // pivot columns of the sweep; DW signals
  __shared__ double s_W[NTRI];        // packed lower triangle of the normal matrix
  __shared__ unsigned long long s_words[kMaxDirWords];
  __shared__ unsigned long long s_vox;
  float* const s_b = reinterpret_cast<float*>(s_h);   // DW signals during set-up (s_h is free then)
  __shared__ int s_flag[NWAVE];

  const int tid = threadIdx.x;
  const int N = P.n_dw;
  const int R = P.n_hr;
  const int Rpad = P.n_hr_pad;
  const int nq = Rpad / NT;   // host guarantees Rpad % NT == 0 and Rpad / 64 <= kMaxDirWords
  const int nwords = nq * NWAVE;          // bitmask words in use (word w: directions 64w..64w+63)
  const int nwio = (R + 63) / 64;         // words per voxel in the caller's mask arrays
  const int wave = tid >> 6;
  const bool row_ok = tid < NSH;

  // statistics of this workgroup (32-bit: well below 2^32 per workgroup); |neg| is summed per lane
  unsigned st_vox = 0, st_solves = 0, st_conv = 0, st_fallback = 0, st_warm_solves = 0;
  unsigned long long st_kl = 0;
  const int lane = tid & 63;
  const bool warm = WARM && P.niter > 0;

  for (;;) {
    unsigned long long vox = 0;
    if constexpr (NT == 64) {
      if (tid == 0) vox = atomicAdd(counter, 1ull);
      vox = uniform64(vox);
    } else {
      wave_sync();
      if (tid == 0) s_vox = atomicAdd(counter, 1ull);
      wave_sync();
      vox = uniform64(s_vox);
    }
    if (vox >= nvox) break;

    // ---- gather the selected DW signals (dwi2fod.cpp:156-173) ----
    bool bad = valid && !valid[vox];
    const float* src = dw + vox * (size_t)ld_dw;
    for (int r = opaque_lane(); r < N; r += NT) {
      const float v = src[P.dwis[r]];
      if (!isfinite(v)) bad = true;
      s_b[r] = v < 0.0f ? 0.0f : v;
    }
    bool any_bad;
    if constexpr (NT == 64) {
      any_bad = __ballot(bad) != 0ull;
    } else {
      const unsigned long long bb = __ballot(bad);
      if ((tid & 63) == 0) s_flag[wave] = bb != 0ull;
      wave_sync();
      any_bad = false;
#pragma unroll
      for (int w = 0; w < NWAVE; ++w) any_bad = any_bad || s_flag[w];
    }
    if (any_bad) {
      if (row_ok) fod[vox * NSH + tid] = 0.0f;
      if (tid == 0) {
        if (iters_out) iters_out[vox] = 0xFFFF;
        if (nneg_out) nneg_out[vox] = 0;
      }
      if (negmask && tid < nwio) negmask[vox * nwio + tid] = 0ull;
      continue;
    }
    wave_sync();

    // ---- CSD::set: Mt_b and the initial F (csd.h:236-243) ----
    // Chunks of kPipe volumes: all loads of a chunk are issued before its FMAs (volumes past N
    // are clamped reads with a zero weight, which leaves the sums unchanged).
    double mtb = 0.0, f_plain = 0.0;
    {
      const int ni = P.n_init;
      const int t = opaque_lane();
      const int rt = t < NSH ? t : 0;
      const int it = t < ni ? t : 0;
      for (int r0 = 0; r0 < N; r0 += kPipe) {
        double m[kPipe], g[kPipe], bv[kPipe];
#pragma unroll
        for (int i = 0; i < kPipe; ++i) {
          const int r = r0 + i < N ? r0 + i : N - 1;
          m[i] = P.M[r * NSH + rt];
          if constexpr (!WARM) g[i] = P.rconvT[r * ni + it];
          bv[i] = r0 + i < N ? (double)s_b[r] : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < kPipe; ++i) {
          mtb = fma(m[i], bv[i], mtb);
          if constexpr (!WARM) f_plain = fma(g[i], bv[i], f_plain);
        }
      }
      if (!row_ok) mtb = 0.0;
      if (tid >= ni) f_plain = 0.0;
    }
    // rconv b for the plain start of a warm launch, from the voxel's row in HBM (the signal
    // buffer in LDS has been reused by then); same order of accumulation as above
    auto plain_start = [&]() __attribute__((always_inline)) -> double {
      const int ni = P.n_init;
      const int t = opaque_lane();
      const int it = t < ni ? t : 0;
      double acc = 0.0;
      for (int r0 = 0; r0 < N; r0 += kPipe) {
        double g[kPipe], bv[kPipe];
#pragma unroll
        for (int i = 0; i < kPipe; ++i) {
          const int r = r0 + i < N ? r0 + i : N - 1;
          g[i] = P.rconvT[r * ni + it];
          const float v = src[P.dwis[r]];
          bv[i] = r0 + i < N ? (double)(v < 0.0f ? 0.0f : v) : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < kPipe; ++i) acc = fma(g[i], bv[i], acc);
      }
      return t < ni ? acc : 0.0;
    };

    // The negative sets as bitmasks held one 64-bit word per lane (lane w: directions 64w..64w+63),
    // so they occupy one VGPR pair instead of scalar registers; |neg| is counted per lane.
    unsigned long long my_old = 0ull;
    int n = 0;
    int last_kl = 0;
    bool converged = false;
    double f = 0.0;
    // pass 0: warm start (proposed F0 and/or negative set), at most warm_budget iterate() calls;
    // pass 1: the plain initialisation with the full niter budget (the reference trajectory)
    for (int pass = warm ? 0 : 1; pass < 2 && !converged; ++pass) {
      const bool proposal = pass == 0 && warm_neg != nullptr;
      const int limit = pass == 0 ? warm_budget : P.niter;
      if constexpr (WARM) {
        const int t = opaque_lane();
        f = pass == 0 ? ((warm_F0 && t < NSH) ? (double)warm_F0[vox * NSH + t] : 0.0) : plain_start();
      } else {
        f = f_plain;
      }
      {
        double A0[NSH];
        const int t = opaque_lane();
        const double* mrow = P.MtM + (t < NSH ? t : 0) * NSH;
#pragma unroll
        for (int j = 0; j < NSH; ++j) A0[j] = mrow[j];
        wave_sync();
        store_lower_row<NSH>(s_W, s_F, A0);
      }
      // the first update may read these rows without an amplitude pass (and its barriers) in
      // between, when a proposed set replaces it: rows written by other wavefronts must be visible
      wave_sync();
      my_old = 0ull;
      bool have_old = false;   // old_neg = [-1] (csd.h:240)
      last_kl = 0;

      for (n = 0; n < limit; ++n) {
        unsigned long long my_new = 0ull;
        bool same = false;
        if (proposal && n == 0) {
          // the proposed negative set replaces the first amplitude evaluation; only bits of
          // existing directions are honoured
          const int l = opaque_lane() & 63;
          if (l < nwio) {
            const int left = R - 64 * l;
            const unsigned long long valid_bits = left >= 64 ? ~0ull : (1ull << left) - 1ull;
            my_new = warm_neg[vox * nwio + l] & valid_bits;
          }
        } else {
          // ---- amplitudes and the negative set (csd.h:246-250) ----
          // One block of NT directions at a time (thread t: direction b*NT + t): NSH independent
          // loads per block, accumulated in coefficient order, then one ballot per wavefront.
          s_F[opaque_lane()] = f;
          wave_sync();
          for (int b = 0; b < nq; ++b) {
            const double* hb = P.HRt + (size_t)b * (NSH * NT) + opaque_lane();
            double a = 0.0;
            double hbuf[2][kPipe], fbuf[2][kPipe];
#pragma unroll
            for (int i = 0; i < kPipe; ++i) {
              if (i < NSH) {
                hbuf[0][i] = hb[i * NT];
                fbuf[0][i] = s_F[i];
              }
            }
#pragma unroll
            for (int kc = 0; kc * kPipe < NSH; ++kc) {
              const int j = kc * kPipe;
              if (j + kPipe < NSH) {
#pragma unroll
                for (int i = 0; i < kPipe; ++i) {
                  if (j + kPipe + i < NSH) {
                    hbuf[(kc + 1) & 1][i] = hb[(j + kPipe + i) * NT];
                    fbuf[(kc + 1) & 1][i] = s_F[j + kPipe + i];
                  }
                }
              }
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int i = 0; i < kPipe; ++i)
                if (j + i < NSH) a = fma(hbuf[kc & 1][i], fbuf[kc & 1][i], a);
              __builtin_amdgcn_sched_barrier(0);
            }
            const int r = opaque_lane() + b * NT;
            const unsigned long long ball = __ballot(r < R && a < P.thr);
            if ((tid & 63) == 0) s_words[b * NWAVE + wave] = ball;
          }
          wave_sync();
          {
            const int l = opaque_lane() & 63;
            my_new = l < nwords ? s_words[l] : 0ull;
          }
          same = have_old && __ballot(my_new != my_old) == 0ull;
        }
        if (same) { converged = true; break; }   // (csd.h:252-253)
        // ---- W = Mt_M + H_neg^T H_neg (csd.h:255-262), updated by the change of the neg set ----
        // Directions that entered (+) or left (-) neg since the last solve are applied to the
        // packed lower triangle in LDS as a signed rank-4 update per MFMA (mfma_update).
        {
          // word w of the change (scanned in order): delta = new ^ old (new alone at the first
          // solve), cur_nw = the new word, for the sign
          int w = 0;
          unsigned long long cur_nw = readlane64(my_new, 0);
          unsigned long long delta = have_old ? (cur_nw ^ readlane64(my_old, 0)) : cur_nw;
          // next direction of the change (entered: +1, left: -1); 0 when none is left
          auto take1 = [&](int& cnt, double& sgn) __attribute__((always_inline)) -> int {
            while (delta == 0ull && w + 1 < nwords) {
              ++w;
              cur_nw = readlane64(my_new, w);
              delta = have_old ? (cur_nw ^ readlane64(my_old, w)) : cur_nw;
            }
            sgn = 0.0;
            if (delta == 0ull) return 0;
            const int bit = __ffsll((long long)delta) - 1;
            delta &= delta - 1;
            sgn = ((cur_nw >> bit) & 1ull) ? 1.0 : -1.0;
            ++cnt;
            return w * 64 + bit;
          };
          auto take4 = [&](int& r0, int& r1, int& r2, int& r3, double& s0, double& s1, double& s2,
                           double& s3) __attribute__((always_inline)) -> int {
            int cnt = 0;
            r0 = take1(cnt, s0);
            r1 = take1(cnt, s1);
            r2 = take1(cnt, s2);
            r3 = take1(cnt, s3);
            return cnt;
          };
          if constexpr (NWAVE == 1) {
            mfma_update<NSH, 1, 0>(s_W, s_F, P.HR, take4);
          } else {
            if (wave == 0) mfma_update<NSH, NWAVE, 0>(s_W, s_F, P.HR, take4);
            else mfma_update<NSH, NWAVE, 1>(s_W, s_F, P.HR, take4);
          }
        }
        wave_sync();
        double A[NSH];
        load_sym_row<NSH>(s_W, A);
        my_old = my_new;
        have_old = true;
        last_kl = __popcll(my_new);
        ++st_solves;
        st_kl += (unsigned long long)last_kl;

        // ---- Gauss-Jordan sweep on [W | Mt_b] (thread i owns row i) ----
        // Step kk eliminates column kk from every other row.  Pivot row kk equals column kk past
        // the diagonal, so each lane publishes its A[kk] as column kk.  The next column is
        // updated first and published before the rest of the row (look-ahead), so its LDS
        // round trip overlaps the remaining FMAs.
        double* const col0 = s_h;
        double* const col1 = s_h + kColStride;
        double rhs = mtb;
        wave_sync();   // earlier readers of the pivot columns are done
        {
          const int t = opaque_lane();
          col0[t] = A[0];
          if (NT != 64) col0[NT + t] = rhs;
        }
        wave_sync();
        double di = 1.0;
        gj_sweep<NSH, NT>(A, rhs, di, col0, col1, tid, std::make_integer_sequence<int, NSH>{});
        f = row_ok ? rhs / di : 0.0;
        if (pass == 0) ++st_warm_solves;
      }
      if (!converged && pass == 0) ++st_fallback;
      if (P.niter == 0 && pass == 1) converged = true;   // -niter 0: the linear deconvolution
    }

    {
      const int t = opaque_lane();
      if (t < NSH) fod[vox * NSH + t] = (float)f;
    }
    if (nneg_out) {
      const unsigned long long last_k = wave_sum((unsigned long long)last_kl);
      if (tid == 0) nneg_out[vox] = (uint16_t)last_k;
    }
    if (tid == 0 && iters_out) iters_out[vox] = (uint16_t)n;
    if (negmask) {
      const int t = opaque_lane();
      if (t < nwio) negmask[vox * nwio + t] = my_old;
    }
    ++st_vox;
    if (converged && P.niter) ++st_conv;
    wave_sync();
  }

  if (stats && wave == 0) {
    const unsigned long long st_k = wave_sum(st_kl);
    if (tid == 0) {
      atomicAdd(stats + 0, (unsigned long long)st_vox);
      atomicAdd(stats + 1, (unsigned long long)st_solves);
      atomicAdd(stats + 2, st_k);
      atomicAdd(stats + 3, (unsigned long long)st_conv);
      atomicAdd(stats + 4, (unsigned long long)st_fallback);
      atomicAdd(stats + 5, (unsigned long long)st_warm_solves);
    }
  }
}

__global__ void reset_counter_kernel(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

template <int NSH, int NT>
hipError_t launch_impl(const CsdDeviceProblem& P, const CsdLaunch& L, hipStream_t stream) {
  auto kern = (L.warm_F0 || L.warm_neg) ? csd_persistent_kernel<NSH, NT, true> : csd_persistent_kernel<NSH, NT, false>;
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
                                                     L.nneg, L.stats, L.warm_F0, L.warm_neg, L.warm_budget,
                                                     L.negmask, L.counter);
  return hipGetLastError();
}

}  // namespace

hipError_t csd_launch(const CsdDeviceProblem& P, const CsdLaunch& L, hipStream_t stream) {
  switch (P.n_sh) {
    case 6:  return launch_impl<6, 64>(P, L, stream);
    case 15: return launch_impl<15, 64>(P, L, stream);
    case 28: return launch_impl<28, 64>(P, L, stream);
    case 45: return launch_impl<45, 64>(P, L, stream);
    case 66: return launch_impl<66, 128>(P, L, stream);
    case 91: return launch_impl<91, 128>(P, L, stream);
    default: return hipErrorInvalidValue;
  }
}

namespace {
template <int NSH, int NT>
hipError_t full_grid_impl(bool warm, int* groups) {
  auto kern = warm ? csd_persistent_kernel<NSH, NT, true> : csd_persistent_kernel<NSH, NT, false>;
  int per_cu = 0, dev = 0, ncu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  *groups = ncu * (per_cu < 1 ? 1 : per_cu);
  return hipSuccess;
}
}  // namespace

hipError_t csd_full_grid(int n_sh, bool warm, int* groups) {
  switch (n_sh) {
    case 6:  return full_grid_impl<6, 64>(warm, groups);
    case 15: return full_grid_impl<15, 64>(warm, groups);
    case 28: return full_grid_impl<28, 64>(warm, groups);
    case 45: return full_grid_impl<45, 64>(warm, groups);
    case 66: return full_grid_impl<66, 128>(warm, groups);
    case 91: return full_grid_impl<91, 128>(warm, groups);
    default: return hipErrorInvalidValue;
  }
}

bool csd_supported_nsh(int n_sh) {
  return n_sh == 6 || n_sh == 15 || n_sh == 28 || n_sh == 45 || n_sh == 66 || n_sh == 91;
}

int csd_threads_for(int n_sh) { return n_sh <= 64 ? 64 : 128; }

}  // namespace mrcsd
