/*
 * mrpeaks — the peaks of the SH series in every voxel on the GPU, as sh2peaks finds them (part of
 * libmrcsd.so; status codes and mrcsd_last_error() as in mrcsd.h).
 *
 * Replaces, for a batch of voxels:
 *   - sh2peaks' Processor (cmd/sh2peaks.cpp:145-300): the search from every direction, the
 *     de-duplication (|dot| > 0.99), the threshold and the choice of the reported peaks
 *   - Math::SH::get_peak and Math::SH::derivatives (core/math/SH.h:514-647)
 * computed in single precision as the reference computes them, or (-fast) from the reference's
 * lookup table of Legendre values (Math::SH::PrecomputedAL, SH.h:402-470).
 */
#ifndef MRPEAKS_H
#define MRPEAKS_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrpeaks_ctx mrpeaks_ctx;

/* The Processor's configuration (sh2peaks.cpp:326-349, 160-170):
 *   n_sh        coefficients per voxel (the SH image's 4th dimension); lmax = LforN(n_sh)
 *               (SH.h:70-73), an even value from 2 to 12
 *   npeaks      -num (default 3); ignored when directions are given
 *   threshold   -threshold (use -INFINITY for none)
 *   seeds       [n_seeds][2] (azimuth, elevation) in radians (-seeds), or NULL for the built-in
 *               60 directions (sh2peaks.cpp:363-424); at most 512
 *   directions  [n_directions][2] (phi, theta) in radians (-direction, converted from degrees by
 *               the caller as sh2peaks.cpp:331 does), or NULL; reports, for each, the peak most
 *               nearly parallel to it
 *   fast        -fast: nonzero takes the associated Legendre values from a table of 512 elevations
 *               (n pi / 511, built at create) by linear interpolation, as PrecomputedAL does
 *               (SH.h:418-470); 0 (the default) evaluates them exactly */
typedef struct mrpeaks_options {
  int n_sh;
  int npeaks;
  float threshold;
  const float* seeds;
  int n_seeds;
  const float* directions;
  int n_directions;
  int fast;
} mrpeaks_options;

void mrpeaks_options_default(mrpeaks_options* opt);
int mrpeaks_create(const mrpeaks_options* opt, mrpeaks_ctx** out);
void mrpeaks_destroy(mrpeaks_ctx* ctx);
/* The number of peaks reported per voxel (3 values each in the output). */
int mrpeaks_npeaks(const mrpeaks_ctx* ctx);

/* Processor::operator() for nvox voxels.  Device pointers:
 *   sh        [nvox][ld_sh] float, the first n_sh values of a row
 *   mask      [nvox] uint8 or NULL: 0 -> an all-NaN voxel (the DataLoader's NaN fill)
 *   peaks_in  [nvox][ld_peaks_in] or NULL (-peaks): the first 3 npeaks values of a row are the
 *             directions to match; a NaN first value gives an all-NaN voxel
 *   out       [nvox][3 npeaks]: amplitude * unit direction of each reported peak, NaN past the last
 *   stats     uint64[4] or NULL, accumulated: voxels searched, Newton iterations, searches that did
 *             not converge in 50 iterations, voxels with more than 64 distinct peaks
 * A peaks_in direction with no accepted peak at |dot| > 0 is reported as NaN (the reference
 * reports whatever its per-thread buffer last held there).
 * The context keeps a device work buffer of 16 B per (voxel, search direction), grown on demand up
 * to 1 GiB (64M searches); larger calls run in voxel passes.  Calls on one context are ordered on
 * the stream they are given; concurrent calls on one context from different streams race. */
int mrpeaks_run(mrpeaks_ctx* ctx, const float* sh, int ld_sh, const uint8_t* mask, const float* peaks_in,
                int ld_peaks_in, size_t nvox, float* out, uint64_t* stats, void* stream);
/* The same with host pointers (synchronous).  Fails with MRCSD_ERR_SETUP (outputs still written)
 * when a voxel has more than 64 distinct accepted peaks: the selection then sees the first 64 in
 * seed order, where the reference chooses among all of them.  The built-in 60 search directions
 * cannot reach this; -seeds with more than 64 directions can. */
int mrpeaks_run_host(mrpeaks_ctx* ctx, const float* sh, int ld_sh, const uint8_t* mask, const float* peaks_in,
                     int ld_peaks_in, size_t nvox, float* out);

#ifdef __cplusplus
}
#endif

#endif
