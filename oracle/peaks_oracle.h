/* CPU restatement of sh2peaks (cmd/sh2peaks.cpp) and Math::SH::get_peak (core/math/SH.h) —
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU peak finder and its benchmark's CPU baseline. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#define ORC_PEAKS_MAX_LMAX 14
#define ORC_PEAKS_MAX_SEEDS 512
#define ORC_PEAKS_FAST_NDIR 512   /* PrecomputedAL's default direction count (SH.h:406) */

void  orc_plm_sph_f(float* array, int lmax, int m, float x, float sin_x);
void  orc_sh_derivatives_f(const float* sh, int lmax, float elevation, float azimuth, float* out6);
float orc_sh_get_peak_f(const float* sh, int lmax, float* dir, int* iters);
int   orc_sh2peaks_voxel(const float* sh, int nsh, int lmax, const float* seeds, int nseeds, int npeaks,
                         float threshold, const float* match, float* out, int* total_iters);
long long orc_sh2peaks_run(const float* sh, int ld, int nsh, size_t nvox, int lmax, const float* seeds, int nseeds,
                           int npeaks, float threshold, const float* match, int ld_match, const uint8_t* mask,
                           int nthreads, float* out);
/* -fast: the lookup table ([ORC_PEAKS_FAST_NDIR][(lmax/2+1)^2], returns the row length) and the run
 * with it (table NULL = the exact recursion, as orc_sh2peaks_run) */
int   orc_peaks_fast_table(int lmax, float* table);
long long orc_sh2peaks_run_fast(const float* sh, int ld, int nsh, size_t nvox, int lmax, const float* seeds, int nseeds,
                                int npeaks, float threshold, const float* match, int ld_match, const uint8_t* mask,
                                int nthreads, const float* table, float* out);
