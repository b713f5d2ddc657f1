/*
 * mrfixel — FOD segmentation into fixels on the GPU, as fod2fixel does it (part of libmrcsd.so;
 * status codes and mrcsd_last_error() as in mrcsd.h).
 *
 * Replaces, for a batch of voxels:
 *   - DWI::FMLS::Segmenter and its operator() (src/dwi/fmls.cpp:144-377): amplitudes on the
 *     segmentation directions, lobe growth in decreasing amplitude order with merges and
 *     retrospective assignments, integral and peak thresholds, Newton revision of the peaks
 *   - DWI::FMLS::IntegrationWeights (fmls.cpp:95-141) and DWI::Directions::FastLookupSet
 *     (src/dwi/directions/set.cpp:336-467), built at create
 *   - fod2fixel's Segmented_FOD_receiver (cmd/fod2fixel.cpp:108-283): per voxel at most -maxnum
 *     fixels, direction from the mean or (-dirpeak) the maximal peak, AFD integral and peak
 *     amplitude, laid out in voxel order with the index image's count and offset
 * in double precision as the reference computes them (results are rounded to float on output, as
 * fod2fixel stores them).
 */
#ifndef MRFIXEL_H
#define MRFIXEL_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrfixel_ctx mrfixel_ctx;

/* The Segmenter's configuration (fmls.cpp:144-158, fmls.h:31-33, load_fmls_thresholds at
 * fmls.cpp:56-91, fod2fixel.cpp:288-330):
 *   n_sh                  coefficients per voxel (the FOD image's 4th dimension); lmax =
 *                         LforN(n_sh), an even value from 2 to 12
 *   integral_threshold    -fmls_integral (default 0); -fmls_no_thresholds sets both thresholds to 0
 *   peak_value_threshold  -fmls_peak_value (default 0.1)
 *   lobe_merge_ratio      -fmls_lobe_merge_ratio (default 1.0: lobes never merge)
 *   maxnum                -maxnum (0, the default: every fixel)
 *   dir_from_peak         -dirpeak: the fixel direction is the lobe's maximal peak instead of its
 *                         weighted mean direction
 *   directions            [n_directions][2] azimuth / elevation in radians of the segmentation
 *                         directions, or NULL for the 1281-direction tessellation fod2fixel uses
 *                         (fod2fixel.cpp:328); 64 to 1536 directions */
typedef struct mrfixel_options {
  int n_sh;
  double integral_threshold;
  double peak_value_threshold;
  double lobe_merge_ratio;
  int maxnum;
  int dir_from_peak;
  const double* directions;
  int n_directions;
} mrfixel_options;

void mrfixel_options_default(mrfixel_options* opt);
int mrfixel_create(const mrfixel_options* opt, mrfixel_ctx** out);
void mrfixel_destroy(mrfixel_ctx* ctx);

/* Segments nvox voxels (one GPU lane per voxel).  Device pointers:
 *   fod       [nvox][ld_fod] float, the first n_sh values of a row are the SH coefficients
 *   mask      [nvox] uint8 or NULL (-mask): 0 -> no fixels
 *   count     [nvox] uint32: fixels of each voxel (the index image's first volume)
 *   offset    [nvox] uint32: the voxel's first fixel, counting voxels in the given order; 0 for a
 *             voxel without fixels (the index image's second volume)
 *   fixels    [capacity][5] float: direction x, y, z, integral (AFD), maximal peak amplitude
 *             (directions.mif, afd, peak_amp; dispersion is integral / peak)
 *   summary   [nvox][5] float or NULL: peak amplitude of fixels 0 and 1 (0 where absent) and the
 *             direction of fixel 0 (NaN where absent)
 *   total     uint64 (device): the number of fixels; when it exceeds capacity, count and offset
 *             are complete and only the fixels below capacity are written: call again with
 *             capacity >= total for the rest
 * Work arrays are kept in the context, for up to 131072 voxels per pass (fewer when a call has
 * fewer, or when half the free device memory holds fewer): about 110 KB per voxel at 1281
 * directions (amplitudes, labels, visiting order and lobe tables sized for as many lobes as
 * directions), 14 GB for a full pass.  Calls on one context are
 * ordered on the stream they are given; concurrent calls on one context race. */
int mrfixel_segment(mrfixel_ctx* ctx, const float* fod, int ld_fod, const uint8_t* mask, size_t nvox, uint32_t* count,
                    uint32_t* offset, float* fixels, size_t capacity, float* summary, uint64_t* total, void* stream);

/* The same with host pointers (synchronous); *total is a host value. */
int mrfixel_segment_host(mrfixel_ctx* ctx, const float* fod, int ld_fod, const uint8_t* mask, size_t nvox,
                         uint32_t* count, uint32_t* offset, float* fixels, size_t capacity, float* summary,
                         uint64_t* total);

/* Host only (no device needed): the tables a segmenter is built on, for directions az_el
 * [n][2] (NULL: the 1281-direction tessellation, n ignored).  Call with NULL arrays to get the
 * sizes in info, then with arrays of those sizes (any may be NULL):
 *   unit        [ndir][3]
 *   adj         [ndir][8]: neighbours on the convex hull of the directions and their antipodes,
 *               ascending, -1 padded (Set::initialise_adjacency, set.cpp:70-313)
 *   weights     [ndir] integration weights
 *   cell_start  [n_az n_el + 1], cell_dirs [n_cell_dirs]: the FastLookupSet grid */
typedef struct mrfixel_set_info {
  int ndir, n_az, n_el, n_cell_dirs;
} mrfixel_set_info;
int mrfixel_direction_set(const double* az_el, int n, mrfixel_set_info* info, double* unit, int* adj, double* weights,
                          int* cell_start, int* cell_dirs);

/* Host only: the PrecomputedAL<double> table (core/math/SH.h:418-434) the peak search
 * interpolates: [nrows][(lmax/2+1)^2], row k at elevation k pi / (nrows - 1). */
int mrfixel_legendre_table(int lmax, int nrows, double* out);

#ifdef __cplusplus
}
#endif

#endif
