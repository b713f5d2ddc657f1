// amp2response's accumulation on gfx950 (cmd/amp2response.cpp:152-191): for every voxel of the
// mask, the shell's gradient directions are expressed in the frame whose z axis is the voxel's
// fibre direction, their elevations (folded into [0, pi/2]) give the zonal-harmonic amplitude
// transform T (core/math/ZSH.h:66-77), and M += T^T T, b += T^T a over the voxels.
//
// Only the elevation enters a zonal harmonic, so the rotation reduces to cos(el) = |d . g| / |g|
// for the unit fibre direction d and gradient direction g: the reference's randomly chosen x and y
// axes of the frame (amp2response.cpp:72-90) change nothing but rounding.  FP64 throughout; each
// thread sums its voxels, a workgroup reduces its threads in a fixed tree, and a second kernel adds
// the workgroups' sums in a fixed order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>
#include <math.h>

#include "resp.h"

namespace mrcsd {
namespace {

constexpr int kThreads = 256;
constexpr double kY00 = 0.28209479177387814;   // 1 / sqrt(4 pi)

// a / f for a divisor known at compile time: y = 1 / f folds to the correctly rounded reciprocal
// and one product with two fused corrections gives the correctly rounded quotient (Markstein's
// theorem; checked against the division on random significands by tests/native/div_const_check.c)
__device__ __forceinline__ double div_const(double a, double f) {
  const double y = 1.0 / f;
  const double q0 = a * y;
  const double r = fma(f, q0, -a);
  return fma(-r, y, q0);
}

// core/math/legendre.h:100-123 for m = 0: the orthonormal P_l^0(x) for even l <= LMAX
template <int LMAX>
__device__ __forceinline__ void zonal_row(double x, double (&t)[LMAX / 2 + 1]) {
  double p[LMAX + 2];
  p[0] = kY00;
  t[0] = kY00;
  if constexpr (LMAX > 0) {
    double f = sqrt(3.0);
    p[1] = x * f * p[0];
#pragma unroll
    for (int l = 2; l <= LMAX; ++l) {
      const double v = x * p[l - 1] - div_const(p[l - 2], f);
      f = sqrt((double)(4 * l * l - 1) / (double)(l * l));
      p[l] = v * f;
    }
#pragma unroll
    for (int l = 2; l <= LMAX; l += 2) t[l / 2] = p[l];
  }
}

template <int LMAX>
__global__ __launch_bounds__(kThreads) void resp_accumulate_kernel(RespLaunch L) {
  constexpr int NZ = LMAX / 2 + 1, NM = NZ * (NZ + 1) / 2, NV = NM + NZ + 1;
  __shared__ double s_red[kThreads];
  double M[NM], b[NZ], cnt = 0.0;
#pragma unroll
  for (int q = 0; q < NM; ++q) M[q] = 0.0;
#pragma unroll
  for (int q = 0; q < NZ; ++q) b[q] = 0.0;
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t v = (size_t)blockIdx.x * kThreads + threadIdx.x; v < L.nvox; v += stride) {
    if (L.mask && !L.mask[v]) continue;
    cnt += 1.0;
    const float* fd = L.fdir + v * (size_t)L.ld_fdir;
    double dx = fd[0], dy = fd[1], dz = fd[2];
    const double n2 = dx * dx + dy * dy + dz * dz;
    if (n2 > 0.0) {   // Eigen's normalize() leaves a zero vector as it is
      const double nrm = sqrt(n2);
      dx /= nrm;
      dy /= nrm;
      dz /= nrm;
    }
    const float* a = L.amp + v * (size_t)L.ld_amp;
    for (int i = 0; i < L.nvol; ++i) {
      const double gx = L.dirs[3 * i], gy = L.dirs[3 * i + 1], gz = L.dirs[3 * i + 2];
      // |R g| is |g| for a unit fibre direction; a zero one leaves R zero (Eigen's normalize keeps a
      // zero vector), so the reference's elevation is acos(0 / 0) and so is this cosine
      const double x = n2 > 0.0 ? fabs(dx * gx + dy * gy + dz * gz) * L.inv_norm[i] : NAN;   // cos of the folded elevation
      double t[NZ];
      zonal_row<LMAX>(x, t);
      const double ai = (double)a[L.vols[i]];
      int q = 0;
#pragma unroll
      for (int r0 = 0; r0 < NZ; ++r0) {
        b[r0] = fma(t[r0], ai, b[r0]);
#pragma unroll
        for (int c0 = 0; c0 <= r0; ++c0) {
          M[q] = fma(t[r0], t[c0], M[q]);
          ++q;
        }
      }
    }
  }
  // fixed-order tree reduction of each value over the workgroup
  double* dst = L.partial + (size_t)blockIdx.x * NV;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double val = q < NM ? M[q] : (q < NM + NZ ? b[q - NM] : cnt);
    s_red[threadIdx.x] = val;
    __syncthreads();
#pragma unroll
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) s_red[threadIdx.x] += s_red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dst[q] = s_red[0];
    __syncthreads();
  }
}

// out[q] = the workgroups' sums of value q (one workgroup per value): thread t adds partials t,
// t + 256, ... in order, then a fixed tree; the same order on every run
__global__ __launch_bounds__(kThreads) void resp_sum_kernel(const double* partial, int nblocks, int nv, double* out) {
  __shared__ double s_red[kThreads];
  const int q = (int)blockIdx.x;
  double s = 0.0;
  for (int k = (int)threadIdx.x; k < nblocks; k += kThreads) s += partial[(size_t)k * nv + q];
  s_red[threadIdx.x] = s;
  __syncthreads();
#pragma unroll
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s_red[threadIdx.x] += s_red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = s_red[0];
}

template <int LMAX>
hipError_t launch(const RespLaunch& L, hipStream_t stream) {
  resp_accumulate_kernel<LMAX><<<dim3((unsigned)L.nblocks), dim3(kThreads), 0, stream>>>(L);
  resp_sum_kernel<<<dim3((unsigned)resp_values(LMAX)), dim3(kThreads), 0, stream>>>(L.partial, L.nblocks, resp_values(LMAX), L.out);
  return hipGetLastError();
}

}  // namespace

hipError_t resp_accumulate_launch(const RespLaunch& L, hipStream_t stream) {
  if (L.nblocks < 1 || L.nvol < 1 || resp_values(L.lmax) > 64) return hipErrorInvalidValue;
  switch (L.lmax) {
    case 0: return launch<0>(L, stream);
    case 2: return launch<2>(L, stream);
    case 4: return launch<4>(L, stream);
    case 6: return launch<6>(L, stream);
    case 8: return launch<8>(L, stream);
    case 10: return launch<10>(L, stream);
    case 12: return launch<12>(L, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mrcsd
