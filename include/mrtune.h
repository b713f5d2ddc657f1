/*
 * mrtune — launch-shape autotuning and captured solve steps (part of libmrcsd.so; conventions as
 * in mrcsd.h).  None of this has a counterpart in MRtrix3, which solves voxel by voxel on CPU
 * threads; it serves callers that solve many volumes of one protocol (dwi2fod on a study,
 * dwi2response's iterations, serving loops):
 *   - the autotuners time candidate launch shapes on the caller's own voxels and keep the fastest
 *     in the context, for its later calls;
 *   - a solve step (optional warm-start inference, then the CSD solve) at fixed device pointers is
 *     captured once as a HIP graph and replayed with one launch.
 */
#ifndef MRTUNE_H
#define MRTUNE_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"
#include "mricls.h"
#include "mrwarm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MRTUNE_MAX_CANDIDATES 8

/* What a tuner tried: candidate values (workgroups, or active-set slots) and their median times. */
typedef struct mrtune_report {
  int n_candidates;
  int values[MRTUNE_MAX_CANDIDATES];
  float ms[MRTUNE_MAX_CANDIDATES];
  int chosen;    /* the value kept */
  int full;      /* CSD: workgroups of the full persistent grid; ICLS: the smallest capacity */
} mrtune_report;

/* Persistent workgroups of mrcsd_run's launches, separately for the plain (warm = 0) and the
 * warm-started (warm = 1) instantiation; 0 = the full grid (CUs x resident workgroups per CU). */
int mrcsd_set_groups(mrcsd_ctx* ctx, int warm, int groups);
int mrcsd_get_groups(const mrcsd_ctx* ctx, int warm);

/* Times mrcsd_run on the given voxels (device pointers as mrcsd_run takes them; `warm` NULL tunes
 * the plain instantiation) with 8/8, 7/8, 6/8, 5/8 and 4/8 of the full grid, median of three runs
 * each, and keeps the fastest for the context.  fod receives the solution.  Synchronous. */
int mrcsd_autotune(mrcsd_ctx* ctx, const float* dw, int ld_dw, const uint8_t* valid, size_t nvox,
                   const mrcsd_warm* warm, float* fod, void* stream, mrtune_report* report);

/* Times mricls_solve on the given right-hand sides with every first-pass capacity from the
 * smallest that holds n + num_eq active constraints up to 128 slots, and keeps the fastest
 * (mricls_set_capacity).  x receives the solution.  Synchronous. */
int mricls_autotune(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox, float* x,
                    void* stream, mrtune_report* report);

typedef struct mrcsd_graph mrcsd_graph;

/* One solve step at fixed device pointers, captured as a HIP graph: with `net`, the fused
 * inference (mrwarm_infer: dwis_dev, b0_dev, n_b0, proposals into negmask) followed by the
 * warm-started mrcsd_run with `budget`; without, the plain mrcsd_run.  The step runs once eagerly
 * (its results are valid on return), then is captured; the context's launch shapes at this time
 * are the ones replayed.  The graph holds no memset node (counters are cleared by a kernel). */
int mrcsd_graph_create(mrcsd_ctx* ctx, mrwarm_net* net, const float* dw, int ld_dw, const uint8_t* valid,
                       size_t nvox, const int* dwis_dev, const int* b0_dev, int n_b0, uint64_t* negmask,
                       int budget, float* fod, const mrcsd_outputs* out, mrcsd_graph** graph);
/* Replays the step on `stream` (asynchronous).  Replays on one graph must not overlap. */
int mrcsd_graph_launch(mrcsd_graph* graph, void* stream);
void mrcsd_graph_destroy(mrcsd_graph* graph);

#ifdef __cplusplus
}
#endif
#endif
