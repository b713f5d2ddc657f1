// Global tractography's external energy (ExternalEnergyComputer::calcEnergy, GT/externalenergy.cpp:
// 245-252) for a batch of voxels: the residual of the track model y = dwi - K t, the per-voxel NNLS
// fit of the isotropic model (the batched ICLS kernel, in double), then the energy
// |y - Ak[:, 1:] f[1:]|^2 / nrows + mu t0 and its sum.
#include <hip/hip_runtime.h>

#include "gt.h"

namespace mrcsd {
namespace {

constexpr int kResVox = 4;       // voxels per workgroup of the residual kernel
constexpr int kMaxCols = 153;    // NforL(16)
constexpr int kEnergyWaves = 4;  // voxels (one per wavefront) per workgroup of the energy kernel

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
