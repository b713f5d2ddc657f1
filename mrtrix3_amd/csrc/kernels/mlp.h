// BF16 MFMA building blocks of the warm-start network (see mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mrcsd {

enum GemmEpilogue { kEpiStore = 0, kEpiBiasGelu = 1, kEpiBias = 2 };

struct GemmArgs {
  const __bf16* A;      // element (m, k) at A[m * sAm + k * sAk]
  long sAm, sAk;
  const __bf16* B;      // element (k, n) at B[k * sBk + n * sBn]
  long sBk, sBn;
  float* C;             // fp32 output (m, n) at C[m * sCm + n * sCn]
  long sCm, sCn;
  __bf16* Cb;           // bf16 activation output (kEpiBiasGelu), row pitch sCbm
  long sCbm;
  const float* bias;    // [N] (kEpiBias, kEpiBiasGelu)
  int M, N, K;
  int epilogue;
  int accumulate;       // kEpiStore: C += result
};

hipError_t gemm_bf16(const GemmArgs& g, hipStream_t stream);

// Fused inference (features, the four layers, proposal) for networks with kInferHidden hidden units
// and at most kInferMaxIn inputs (padded to a multiple of 32).
constexpr int kInferHidden = 256;
constexpr int kInferMaxIn = 128;
struct InferArgs {
  const float* dw;            // [M][ld] DW rows
  int ld;
  size_t M;
  const int* dwis;            // [n_in] input volumes
  int n_in, k0;               // k0 = n_in rounded up to 32
  const int* b0;              // [n_b0] b=0 volumes
  int n_b0;
  const __bf16* WT[4];        // transposed weights [npad][kpad], layer l row pitch ldw[l]
  int ldw[4];
  const float* bias[4];
  int nsh, nlog;
  float* F0;                  // [M][nsh] or null
  unsigned long long* negmask;   // [M][(nlog + 63) / 64]
};
hipError_t infer_propose(const InferArgs& a, hipStream_t stream);
hipError_t transpose_pad(const __bf16* W, int fin, int fout, __bf16* WT, int kpad, int npad, hipStream_t stream);
hipError_t gelu_backward(const float* pre, const float* dact, __bf16* dpre_bf, float* dpre, size_t n,
                         hipStream_t stream);
hipError_t to_bf16(const float* x, __bf16* y, size_t n, hipStream_t stream);
hipError_t colsum(const float* x, int M, int N, float* out, hipStream_t stream);
hipError_t adam_step(float* w, __bf16* wb, const float* g, float* m, float* v, size_t n, float lr, float b1,
                     float b2, float eps, int t, float wd, hipStream_t stream);
hipError_t features(const float* dw, int ld, size_t M, const int* dwis, int n_in, const int* b0, int n_b0,
                    __bf16* x, float* s0, hipStream_t stream);
hipError_t propose(const float* out, const float* s0, size_t M, int nsh, int nlog, float* F0,
                   unsigned long long* negmask, hipStream_t stream);
hipError_t warm_loss(const float* out, const float* tgt_coef, const uint8_t* tgt_neg, int M, int nsh, int nlog,
                     float wc, float wl, float* dout, float* loss, hipStream_t stream);

}  // namespace mrcsd
