/*
 * mrwarm — the learned warm start of the GPU CSD solver (part of libmrcsd.so; conventions as in
 * mrcsd.h).  It has no counterpart in MRtrix3: the reference always starts CSD::iterate from the
 * linear deconvolution F = rconv * b (src/dwi/sdeconv/csd.h:236-243).
 *
 * A small MLP maps each voxel's b=0-normalised DW signals to (a) SH coefficients and (b) one
 * logit per constraint direction (is the direction in the final negative set?).  The solver then
 * starts from the proposed negative set S: it solves with S and accepts the voxel only if the
 * fixed-point test neg(F(S)) == S holds — for threshold 0 this certifies the unique minimiser of
 * the CSD objective, so accepted voxels carry the plain solver's answer.  Voxels that fail within
 * a small budget of solves are re-solved from the plain initialisation (mrcsd_warm.budget).
 *
 * Network: n_in -> hidden -> hidden -> hidden -> (nsh + nlog), GELU, bf16 weights and activations,
 * fp32 accumulation (v_mfma_f32_16x16x32_bf16), fp32 master weights and Adam.
 */
#ifndef MRWARM_H
#define MRWARM_H

#include <stddef.h>
#include <stdint.h>

#include "mrcsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrwarm_net mrwarm_net;

int mrwarm_create(int n_in, int hidden, int nsh, int nlog, uint64_t seed, mrwarm_net** out);
void mrwarm_destroy(mrwarm_net* net);
int mrwarm_dims(const mrwarm_net* net, int* n_in, int* hidden, int* nsh, int* nlog);
/* A network saved by the Python mirror (WarmNet.save: an .npz of `dims` = [n_in, hidden, nsh, nlog]
 * as integers and `params` = the float32 parameters in mrwarm_get_params order), read without
 * executing anything from the file. */
int mrwarm_load(const char* path, mrwarm_net** out);
/* The layer sizes and parameter count such a file holds (no GPU needed). */
int mrwarm_file_dims(const char* path, int* n_in, int* hidden, int* nsh, int* nlog, size_t* n_params);
/* Number of parameters and their order: for each layer, W [fan_in][fan_out] then b [fan_out]. */
size_t mrwarm_num_params(const mrwarm_net* net);
int mrwarm_get_params(const mrwarm_net* net, float* host, size_t n);
int mrwarm_set_params(mrwarm_net* net, const float* host, size_t n);

/* Network inputs from DW rows (device): x[m][i] = dw[m][dwis[i]] / s0[m], s0 = mean of the b=0
 * volumes b0[] (1 when not positive).  x: [M][n_in] bf16 (opaque 2-byte storage), s0: [M] float. */
int mrwarm_features(const float* dw, int ld_dw, size_t M, const int* dwis_dev, int n_in,
                    const int* b0_dev, int n_b0, void* x_bf16, float* s0, void* stream);

/* Forward pass: out [M][nsh + nlog] float (device). */
int mrwarm_forward(mrwarm_net* net, const void* x_bf16, size_t M, float* out, void* stream);

/* One Adam step on a batch: loss = wc * mean ||coef - tgt_coef||^2 + wl * mean BCE(logits, neg).
 * tgt_coef [M][nsh] float, tgt_neg [M][nlog] uint8 (device).  loss_out (host, 2 floats) may be NULL
 * (then nothing is synchronised). */
int mrwarm_train_step(mrwarm_net* net, const void* x_bf16, const float* tgt_coef, const uint8_t* tgt_neg,
                      size_t M, float lr, float wc, float wl, float* loss_out, void* stream);

/* Gradients of that loss w.r.t. all parameters (host, mrwarm_num_params floats; synchronous). */
int mrwarm_gradients(mrwarm_net* net, const void* x_bf16, const float* tgt_coef, const uint8_t* tgt_neg,
                     size_t M, float wc, float wl, float* grads_host, float* loss_out);

/* Proposal for the solver from the forward output: F0[m][c] = out[m][c] * s0[m] (c < nsh) and
 * negmask[m][w] bit r set when logit r > 0 (words = (nlog + 63) / 64). */
int mrwarm_propose(const float* out, const float* s0, size_t M, int nsh, int nlog, float* F0,
                   uint64_t* negmask, void* stream);

/* Features, forward pass and proposal fused in one kernel (activations stay in on-chip memory):
 * the same results as mrwarm_features + mrwarm_forward + mrwarm_propose, without the [M][hidden]
 * intermediates.  dw: [M][ld_dw] DW rows (device), dwis_dev [n_in] and b0_dev [n_b0] volume indices
 * (device); F0 [M][nsh] (may be NULL), negmask [M][(nlog + 63) / 64] (device).  Needs hidden = 256,
 * n_in <= 128 and nsh + nlog <= 512 (MRCSD_ERR_ARG otherwise). */
int mrwarm_infer(mrwarm_net* net, const float* dw, int ld_dw, size_t M, const int* dwis_dev, const int* b0_dev,
                 int n_b0, float* F0, uint64_t* negmask, void* stream);

#ifdef __cplusplus
}
#endif
#endif
