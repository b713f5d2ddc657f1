// Device-side description of FOD segmentation (fod2fixel's FMLS, src/dwi/fmls.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

constexpr int kFmlsMaxDirs = 1536;    // directions of the segmentation set (1281 for fod2fixel)
constexpr int kFmlsMaxLobes = 64;     // positive lobes a voxel's first pass holds in LDS
constexpr int kFmlsMaxLmax = 12;

struct FmlsDeviceSet {
  int ndir;
  const double* unit;        // [ndir][3]
  const int16_t* adj;        // [ndir][8], ascending, -1 padded
  const double* weights;     // [ndir] integration weights
  const double* sht;         // [ndir][nsh] amplitude transform
  int nsh, lmax;
  // nearest-direction grid (FastLookupSet)
  int n_az, n_el;
  double az_step, el_step, az_begin, el_begin;
  const int* cell_start;     // [n_az n_el + 1]
  const int16_t* cell_dirs;
  const double* al;          // [al_ndir][(lmax/2+1)^2] precomputed associated Legendre values
  int al_ndir;
  double integral_threshold, peak_threshold, merge_ratio;
  int maxnum;                // fixels kept per voxel, 0 = all
  int dir_from_peak;         // fixel direction: 1 the maximal peak, 0 the lobe's mean direction
};

struct FmlsLaunch {
  const float* fod;          // [nvox][ld_fod]: the first nsh values of a row are the SH coefficients
  int ld_fod;
  const uint8_t* valid;      // [nvox] or null
  size_t nvox;
  double* amps;              // [chunk][ndir] amplitudes of the current chunk
  size_t chunk;              // voxels per amplitude pass
  uint32_t* count;           // [nvox] fixels of each voxel
  uint32_t* offset;          // [nvox] first fixel of the voxel in `fixels` (voxel order; 0 when none)
  float* fixels;             // [capacity][5]: direction xyz, integral (AFD), maximal peak amplitude
  float* stage;              // [capacity][5]: the same in the order voxels finish
  uint32_t* stage_off;       // [nvox]
  size_t capacity;
  unsigned long long* total; // fixels found in the batch (above capacity: outputs incomplete)
  float* summary;            // [nvox][5] or null: maximal peak amplitude of fixels 0 and 1 (0 when
                             // absent), direction of fixel 0 (NaN when absent)
  unsigned long long* counters;  // [4] work counters
  uint32_t* queue;           // [chunk] voxels of the chunk whose lobes do not fit the LDS tables
  double* big_ws;            // [big_blocks][fmls_big_stride(ndir)] lobe tables of the second pass
  int big_blocks;
  uint32_t* scan_ws;         // [fmls_scan_blocks(nvox) + 1]
};

// Doubles of second-pass workspace per workgroup: lobe peak, mean (3), integral, revised peak
// directions (3), then head, tail and peak count as int16.
__host__ __device__ inline size_t fmls_big_stride(int ndir) { return 8 * (size_t)ndir + (3 * (size_t)ndir + 3) / 4; }
// Voxel-count blocks of the offset scan (2048 voxels each).
inline size_t fmls_scan_blocks(size_t nvox) { return (nvox + 2047) / 2048; }

hipError_t fmls_launch(const FmlsDeviceSet& S, const FmlsLaunch& L, hipStream_t stream);

// Whether the segmentation kernels may be launched.  They are switched off: on their first run on
// an MI355X the launch sequence did not complete, and no cause has been found in the source, the
// generated code or the CPU emulation of the kernel (which completes and matches the oracle).
bool fmls_gpu_enabled();

}  // namespace mrcsd
