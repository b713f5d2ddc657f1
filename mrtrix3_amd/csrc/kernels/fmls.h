// Device-side description of FOD segmentation (fod2fixel's FMLS, src/dwi/fmls.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

constexpr int kFmlsMaxDirs = 1536;    // directions of the segmentation set (1281 for fod2fixel)
constexpr int kFmlsMaxLmax = 12;
constexpr int kFmlsBlock = 64;        // voxels per workgroup of the segmentation and offset kernels

// Row stride (elements) of the voxel-major amplitude array: ndir rounded up to 32 doubles (256 bytes).
__host__ __device__ inline int fmls_amp_ld(int ndir) { return (ndir + 31) & ~31; }

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
  int lanes8, lanes16;       // voxels per workgroup of the growth kernels with 8- and 16-bit labels
};

// Voxels per workgroup of a growth kernel whose LDS labels take `bytes` each (ndir per voxel): as
// many as let two workgroups share a CU's LDS within one workgroup's limit, at most 64.
inline int fmls_grow_lanes(int ndir, int bytes, int lds_cu, int lds_block) {
  const long cu = lds_cu > 0 ? lds_cu : lds_block;
  const long wg = cu / 2 < lds_block ? cu / 2 : lds_block;
  const long n = wg / ((long)ndir * bytes);
  return (int)(n < 64 ? n : 64);
}

// Per-voxel work arrays of one chunk.  The amplitudes are voxel-major (row s of fmls_amp_ld(ndir)
// doubles for the voxel in slot s), read by one wavefront per voxel; the other arrays are
// lane-interleaved: element i of the voxel in slot s lives at [i * chunk + s], so the lanes of a
// wavefront (consecutive slots) touching the same element index touch one contiguous run of
// memory.  Every table is sized for the worst case (as many lobes and peaks as directions), so no
// voxel overflows.
struct FmlsWork {
  double* amps;              // [chunk][fmls_amp_ld(ndir)] amplitudes on the directions
  int16_t* lab;              // [ndir] lobe of each direction; -1 none; -2 - l: waits for lobe l
  uint32_t* ordm;            // [ndir] positive directions in visiting order (bits 0-15) and which of
                             // their neighbours are visited before them (bit 16 + r: neighbour r)
  int16_t* nxt;              // [ndir] peak lists through the seed directions
  double* peak;              // [ndir] lobe records: maximal peak amplitude
  double* mean;              // [3 ndir] amplitude-weighted direction sum, then the fixel direction
  double* integ;             // [ndir]
  int16_t* head;             // [ndir] first and last seed of the lobe's peak list
  int16_t* tail;
  int16_t* np;               // [ndir] peaks of the lobe; -1: lobe dropped by a threshold
  double* rev;               // [3 ndir] the amplitudes in visiting order, then the peak directions
                             // during the revision
  int32_t* npos;             // [1] positive directions to visit (0: the voxel has no positive lobe)
  int32_t* nlobes;           // [1] lobes of the voxel, kept and dropped
  uint32_t* items;           // [ndir x chunk] the chunk's lobes as slot << 11 | lobe, voxel by voxel
  unsigned long long* nitems;   // [1] items of the chunk
};

struct FmlsLaunch {
  const float* fod;          // [nvox][ld_fod]: the first nsh values of a row are the SH coefficients
  int ld_fod;
  const uint8_t* valid;      // [nvox] or null
  size_t nvox;
  size_t chunk;              // voxels per pass (slots of the work arrays), a multiple of kFmlsBlock
  FmlsWork work;
  uint32_t* count;           // [nvox] fixels of each voxel
  uint32_t* offset;          // [nvox] first fixel of the voxel in `fixels` (voxel order; 0 when none)
  float* fixels;             // [capacity][5]: direction xyz, integral (AFD), maximal peak amplitude
  size_t capacity;
  unsigned long long* total; // fixels found in the batch (above capacity: fixels incomplete)
  float* summary;            // [nvox][5] or null: maximal peak amplitude of fixels 0 and 1 (0 when
                             // absent), direction of fixel 0 (NaN when absent)
  unsigned long long* sums;  // [chunk / kFmlsBlock + 1] fixel counts of the blocks of a chunk
};

// Bytes of the work arrays for `chunk` slots over `ndir` directions (each array 256-byte aligned).
size_t fmls_work_bytes(int ndir, size_t chunk);
// The arrays carved out of one allocation of fmls_work_bytes(ndir, chunk) bytes.
FmlsWork fmls_work_carve(void* base, int ndir, size_t chunk);

hipError_t fmls_launch(const FmlsDeviceSet& S, const FmlsLaunch& L, hipStream_t stream);

}  // namespace mrcsd
