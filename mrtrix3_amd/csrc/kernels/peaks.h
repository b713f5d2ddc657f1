// Device-side description of a batched sh2peaks run (cmd/sh2peaks.cpp, Math::SH::get_peak).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

constexpr int kPeaksMaxSeeds = 512;     // search directions per voxel
constexpr int kPeaksMaxAccepted = 64;   // distinct peaks kept per voxel (one per lane)
constexpr int kPeaksMaxOut = 64;        // reported peaks per voxel (3 values each)

struct PeaksLaunch {
  const float* sh;          // [nvox][ld_sh]: the first n_sh values of a row are the coefficients
  int ld_sh, n_sh, lmax;
  size_t nvox;
  const float* seeds;       // [nseeds][3] unit search directions (host-made, as Direction(phi, theta))
  int nseeds;
  int npeaks;
  float threshold;
  const float* match;       // [npeaks][3] (ld_match 0: every voxel) or [nvox][ld_match], or null
  int ld_match;
  const uint8_t* mask;      // [nvox] or null
  float* out;               // [nvox][3 npeaks]
  unsigned long long* stats;    // [4] or null: voxels searched, Newton iterations, failed searches,
                                // voxels whose distinct peaks exceeded kPeaksMaxAccepted
  unsigned long long* counter;  // search work counter
  float4* work;                 // [work_items] search results: amplitude, x, y, z
  size_t work_items;            // capacity of `work`; voxels are processed in chunks that fit
  const float* fast_table;      // -fast: [kPeaksFastRows][(lmax/2+1)^2] Legendre values, or null (exact)
};

// rows of the -fast lookup table: elevations n pi / (kPeaksFastRows - 1) (SH.h:406, 418-434)
constexpr int kPeaksFastRows = 512;

// Results of one (voxel, search direction) pair per work item; the buffer a launch needs for nvox
// voxels at most (the launch processes chunks of work_items / nseeds voxels).
inline size_t peaks_work_items(size_t nvox, int nseeds) {
  const size_t cap = size_t(1) << 26;   // 64M searches = 1 GiB (a 646k-voxel volume at 60 seeds in one pass)
  const size_t all = nvox * size_t(nseeds);
  return all < cap ? all : (cap / size_t(nseeds)) * size_t(nseeds);
}

hipError_t peaks_launch(const PeaksLaunch& L, hipStream_t stream);

}  // namespace mrcsd
