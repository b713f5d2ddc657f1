// Device-side accumulation of amp2response's normal equations (cmd/amp2response.cpp:152-191).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

constexpr int kRespMaxLmax = 12;   // zonal coefficients 0, 2, ..., 12

struct RespLaunch {
  const float* amp;        // [nvox][ld_amp]: DW amplitudes, volume axis contiguous
  int ld_amp;
  const float* fdir;       // [nvox][ld_fdir]: fibre direction in the first three values
  int ld_fdir;
  const uint8_t* mask;     // [nvox] or null (every voxel)
  size_t nvox;
  const double* dirs;      // [nvol][3] the shell's unit gradient directions
  const double* inv_norm;  // [nvol] 1 / |direction|
  const int* vols;         // [nvol] their volume indices
  int nvol, lmax;
  double* partial;         // [nblocks][resp_values(lmax)] per-workgroup sums
  double* out;             // [resp_values(lmax)]: lower triangle of M (row-major), b, voxel count
  int nblocks;
};

inline int resp_nz(int lmax) { return lmax / 2 + 1; }
inline int resp_values(int lmax) { return resp_nz(lmax) * (resp_nz(lmax) + 1) / 2 + resp_nz(lmax) + 1; }

hipError_t resp_accumulate_launch(const RespLaunch& L, hipStream_t stream);

}  // namespace mrcsd
