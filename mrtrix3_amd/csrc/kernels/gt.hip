// Global tractography's external energy (ExternalEnergyComputer::calcEnergy, GT/externalenergy.cpp:
// 242-249) for a batch of voxels: the residual of the track model y = dwi - K t, the per-voxel NNLS
// fit of the isotropic model, then the energy |y - Ak[:, 1:] f[1:]|^2 / nrows + mu t0 and its sum.
//
// Up to 4 NNLS unknowns (WM + 3 isotropic tissues) one fused kernel does it all, a lane per voxel:
// the residual rows are formed 8 at a time from the voxel's TOD and signal (staged in LDS) and only
// g = Ak^T y and |y|^2 are kept; the NNLS minimiser of |Ak f - y|^2 (unique: Ak has full column
// rank) is the one active set of the 2^n whose equality-constrained solution z = (G_SS)^-1 g_S
// (G = Ak^T Ak, the inverses prepared on the host) is feasible (z >= 0) and optimal (the gradient
// g - G z <= 0 off the set), every lane testing the same sets in the same order; the energy is
// (|y|^2 - 2 f'.g' + f'^T G' f') / nrows + mu t0 over the isotropic part ('), in FP64.  Larger
// problems take the residual kernel, the batched ICLS kernel and the energy kernel.
#include <hip/hip_runtime.h>

#include "gt.h"

namespace mrcsd {
namespace {

constexpr int kResVox = 4;       // voxels per workgroup of the residual kernel
constexpr int kMaxCols = 153;    // NforL(16)
constexpr int kEnergyWaves = 4;  // voxels (one per wavefront) per workgroup of the energy kernel
constexpr int kFusedN = 4;       // NNLS unknowns of the fused kernel at most

__global__ __launch_bounds__(256) void gt_residual_kernel(const double* __restrict__ Kt, int nrows, int ncols,
                                                          const float* __restrict__ dwi, int ld_dwi,
                                                          const float* __restrict__ tod, int ld_tod, size_t nvox,
                                                          double* __restrict__ y) {
  __shared__ double ts[kResVox][kMaxCols];
  const size_t v0 = (size_t)blockIdx.x * kResVox;
  for (int i = threadIdx.x; i < kResVox * ncols; i += 256) {
    const int vl = i / ncols, c = i - vl * ncols;
    const size_t v = v0 + vl;
    ts[vl][c] = v < nvox ? (double)tod[v * (size_t)ld_tod + c] : 0.0;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kResVox * nrows; q += 256) {
    const int vl = q / nrows, r = q - vl * nrows;
    const size_t v = v0 + vl;
    if (v >= nvox) continue;
    double kt = 0.0;
    for (int c = 0; c < ncols; ++c) kt = fma(Kt[(size_t)c * nrows + r], ts[vl][c], kt);
    y[v * (size_t)nrows + r] = (double)dwi[v * (size_t)ld_dwi + r] - kt;
  }
}

__global__ __launch_bounds__(64 * kEnergyWaves) void gt_energy_kernel(const double* __restrict__ Ak, int nrows, int nf,
                                                                      const double* __restrict__ y,
                                                                      const double* __restrict__ fk,
                                                                      const float* __restrict__ tod, int ld_tod,
                                                                      double mu, size_t nvox, float* __restrict__ eext,
                                                                      float* __restrict__ fiso,
                                                                      double* __restrict__ partial) {
  __shared__ double se[kEnergyWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t v = (size_t)blockIdx.x * kEnergyWaves + wave;
  double e = 0.0;
  if (v < nvox) {
    const double* f = fk + v * (size_t)(nf + 1);
    double acc = 0.0;
    for (int r = lane; r < nrows; r += 64) {
      double res = y[v * (size_t)nrows + r];
      for (int j = 0; j < nf; ++j) res -= Ak[(size_t)r * (nf + 1) + j + 1] * f[j + 1];
      acc = fma(res, res, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    e = acc / (double)nrows + mu * (double)tod[v * (size_t)ld_tod];
    if (lane == 0) eext[v] = (float)e;
    if (fiso && lane < nf) fiso[v * (size_t)nf + lane] = (float)f[lane + 1];
  }
  if (lane == 0) se[wave] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kEnergyWaves; ++w) s += se[w];
    partial[blockIdx.x] = s;
  }
}

// The fused form (kFusedN unknowns at most, a 64-voxel workgroup).  tab: for each subset s of the
// N unknowns (bit i: unknown i free), Z[s] [N][N] with Z[s][i][j] = (G_SS)^-1 for i, j in s and 0
// elsewhere, then G [N][N].  LDS: the workgroup's TOD tile [ncols][64] (float) and the current
// 8-row block of K [ncols][8] (double, read by every lane at the same address).
template <int N>
__global__ __launch_bounds__(64) void gt_fused_kernel(const double* __restrict__ Kt, int nrows, int ncols,
                                                      const double* __restrict__ Ak, const double* __restrict__ tab,
                                                      const float* __restrict__ dwi, int ld_dwi,
                                                      const float* __restrict__ tod, int ld_tod, double mu,
                                                      size_t nvox, float* __restrict__ eext,
                                                      float* __restrict__ fiso, double* __restrict__ partial) {
  extern __shared__ double s_dyn[];
  double* s_k = s_dyn;                                  // [ncols][8]
  float* s_tod = reinterpret_cast<float*>(s_dyn + (size_t)ncols * 8);   // [ncols][64]
  const int lane = threadIdx.x;
  const size_t v0 = (size_t)blockIdx.x * 64;
  const int nv = nvox - v0 < 64 ? (int)(nvox - v0) : 64;
  for (int i = lane; i < nv * ncols; i += 64) {
    const int vl = i / ncols, c = i - vl * ncols;
    s_tod[c * 64 + vl] = tod[(v0 + vl) * (size_t)ld_tod + c];
  }
  double g[N], yy = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) g[k] = 0.0;
  const float* drow = dwi + (v0 + (size_t)(lane < nv ? lane : 0)) * (size_t)ld_dwi;
  for (int r0 = 0; r0 < nrows; r0 += 8) {
    const int nr = nrows - r0 < 8 ? nrows - r0 : 8;
    __syncthreads();   // the previous block of K is no longer read (and the TOD tile is written)
    for (int i = lane; i < ncols * 8; i += 64) {
      const int c = i >> 3, j = i & 7;
      s_k[i] = j < nr ? Kt[(size_t)c * nrows + r0 + j] : 0.0;
    }
    __syncthreads();
    if (lane < nv) {
      // residual rows r0..r0+7: kt = sum_c K[r][c] t_c by FMA in ascending c, y = dwi - kt
      double acc[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = 0.0;
#pragma unroll 4
      for (int c = 0; c < ncols; ++c) {
        const double t = (double)s_tod[c * 64 + lane];
        const double* kc = s_k + c * 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = fma(kc[j], t, acc[j]);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (j < nr) {
          const double y = (double)drow[r0 + j] - acc[j];
          yy = fma(y, y, yy);
#pragma unroll
          for (int k = 0; k < N; ++k) g[k] = fma(Ak[(size_t)(r0 + j) * N + k], y, g[k]);
        }
      }
    }
  }
  double e = 0.0;
  if (lane < nv) {
    // the NNLS minimiser: the active set whose solution satisfies the optimality conditions (the
    // least violation, scaled by the largest |g|, should rounding leave none exactly)
    const double* G = tab + (1 << N) * N * N;
    double gs = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) gs = fmax(gs, fabs(g[k]));
    gs = gs > 0.0 ? gs : 1.0;
    double best = INFINITY, x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = 0.0;
#pragma unroll
    for (int s = 0; s < (1 << N); ++s) {
      const double* Z = tab + (size_t)s * N * N;
      double z[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) a = fma(Z[i * N + j], g[j], a);
        z[i] = a;
      }
      double viol = -INFINITY;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        if ((s >> i) & 1) {
          viol = fmax(viol, -z[i]);
        } else {
          double w = g[i];
#pragma unroll
          for (int j = 0; j < N; ++j) w = fma(-G[i * N + j], z[j], w);
          viol = fmax(viol, w);
        }
      }
      viol /= gs;
      if (viol < best) {
        best = viol;
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = z[i] > 0.0 ? z[i] : 0.0;
      }
    }
    // |y - A' f'|^2 over the isotropic unknowns 1..N-1
    double q = yy;
#pragma unroll
    for (int j = 1; j < N; ++j) {
      q = fma(-2.0 * x[j], g[j], q);
#pragma unroll
      for (int k = 1; k < N; ++k) q = fma(x[j] * x[k], G[j * N + k], q);
    }
    const size_t v = v0 + lane;
    e = q / (double)nrows + mu * (double)s_tod[lane];
    eext[v] = (float)e;
    if (fiso)
#pragma unroll
      for (int j = 1; j < N; ++j) fiso[v * (size_t)(N - 1) + (j - 1)] = (float)x[j];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if (lane == 0) partial[blockIdx.x] = e;
}

// The workgroups' sums added in a fixed order: a strided sum per thread, then a fixed tree.
__global__ __launch_bounds__(256) void gt_sum_kernel(const double* __restrict__ partial, size_t n,
                                                     double* __restrict__ total) {
  __shared__ double s[256];
  double a = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) a += partial[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s[0];
}

}  // namespace

size_t gt_energy_blocks(size_t nvox) { return (nvox + kEnergyWaves - 1) / kEnergyWaves; }

size_t gt_fused_blocks(size_t nvox) { return (nvox + 63) / 64; }

static size_t gt_fused_lds(int ncols) { return size_t(ncols) * (8 * sizeof(double) + 64 * sizeof(float)); }

bool gt_fused_fits(int nrows, int ncols, int n) {
  return n >= 1 && n <= kFusedN && nrows >= 1 && ncols >= 1 && gt_fused_lds(ncols) <= 65536;
}

hipError_t gt_fused(const double* Kt, int nrows, int ncols, const double* Ak, int n, const double* tab,
                    const float* dwi, int ld_dwi, const float* tod, int ld_tod, double mu, size_t nvox, float* eext,
                    float* fiso, double* partial, size_t partial_cap, double* total, hipStream_t stream) {
  if (!gt_fused_fits(nrows, ncols, n) || ld_dwi < nrows || ld_tod < ncols) return hipErrorInvalidValue;
  const size_t blocks = gt_fused_blocks(nvox);
  if (blocks > partial_cap) return hipErrorInvalidValue;
  const size_t lds = gt_fused_lds(ncols);
  if (blocks) {
    const dim3 grid((unsigned)blocks), block(64);
    switch (n) {
      case 1: gt_fused_kernel<1><<<grid, block, lds, stream>>>(Kt, nrows, ncols, Ak, tab, dwi, ld_dwi, tod, ld_tod, mu, nvox, eext, fiso, partial); break;
      case 2: gt_fused_kernel<2><<<grid, block, lds, stream>>>(Kt, nrows, ncols, Ak, tab, dwi, ld_dwi, tod, ld_tod, mu, nvox, eext, fiso, partial); break;
      case 3: gt_fused_kernel<3><<<grid, block, lds, stream>>>(Kt, nrows, ncols, Ak, tab, dwi, ld_dwi, tod, ld_tod, mu, nvox, eext, fiso, partial); break;
      default: gt_fused_kernel<4><<<grid, block, lds, stream>>>(Kt, nrows, ncols, Ak, tab, dwi, ld_dwi, tod, ld_tod, mu, nvox, eext, fiso, partial); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  gt_sum_kernel<<<1, 256, 0, stream>>>(partial, blocks, total);
  return hipGetLastError();
}

hipError_t gt_residual(const double* Kt, int nrows, int ncols, const float* dwi, int ld_dwi, const float* tod,
                       int ld_tod, size_t nvox, double* y, hipStream_t stream) {
  if (nrows < 1 || ncols < 1 || ncols > kMaxCols || ld_dwi < nrows || ld_tod < ncols) return hipErrorInvalidValue;
  if (nvox == 0) return hipSuccess;
  const size_t blocks = (nvox + kResVox - 1) / kResVox;
  gt_residual_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(Kt, nrows, ncols, dwi, ld_dwi, tod, ld_tod, nvox,
                                                                       y);
  return hipGetLastError();
}

hipError_t gt_energy(const double* Ak, int nrows, int nf, const double* y, const double* fk, const float* tod,
                     int ld_tod, double mu, size_t nvox, float* eext, float* fiso, double* partial,
                     size_t partial_cap, double* total, hipStream_t stream) {
  if (nrows < 1 || nf < 0 || nf > 63 || ld_tod < 1) return hipErrorInvalidValue;
  const size_t blocks = gt_energy_blocks(nvox);
  if (blocks > partial_cap) return hipErrorInvalidValue;
  if (blocks)
    gt_energy_kernel<<<dim3((unsigned)blocks), dim3(64 * kEnergyWaves), 0, stream>>>(Ak, nrows, nf, y, fk, tod, ld_tod,
                                                                                     mu, nvox, eext, fiso, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  gt_sum_kernel<<<1, 256, 0, stream>>>(partial, blocks, total);
  return hipGetLastError();
}

}  // namespace mrcsd
