// BF16 MFMA kernels for the warm-start network (a small MLP that proposes each voxel's SH
// coefficients and negative-constraint set before the CSD solve).
//
//   gemm:   C[m, n] = epi( sum_k A[m, k] B[k, n] )  with bf16 A, B (any strides), fp32 accumulate
//           epilogues: plain fp32 store (optionally accumulating into C), bias + GELU (stores the
//           fp32 pre-activation and the bf16 activation), bias only (fp32)
//   gelu_backward, adam, loss gradient, fp32 <-> bf16 conversion kernels.
//
// GEMM tiling: 256 threads (4 wavefronts) per 64 x 64 output tile; each wavefront owns a 32 x 32
// quadrant as 2 x 2 tiles of v_mfma_f32_16x16x32_bf16.  K advances 32 at a time: the A tile
// (64 x 32) and B tile (32 x 64) are staged through LDS with k contiguous, so any operand layout
// (row-major, transposed) is read with the same code and each lane then reads its 8 consecutive
// k values with one 16-byte LDS load.  Out-of-range rows / columns / k are zero-filled.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"

namespace mrcsd {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int LDS_K = TK + 8;   // row pitch in bf16 (16-byte aligned rows, fewer bank conflicts)

// tanh(u) = sign(u) (1 - e) / (1 + e), e = exp(-2|u|), from the hardware exp2 and reciprocal: about
// 1e-7 absolute error (the activations it feeds are rounded to bf16, 8 significant bits), seven
// instructions instead of a branchy library tanhf.  e underflows to 0 for large |u| (t = +-1).
__device__ __forceinline__ float tanh_fast(float u) {
  const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(u));   // exp(-2|u|) = 2^(-2 log2(e) |u|)
  const float t = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
  return copysignf(t, u);
}

__device__ __forceinline__ float gelu_f(float x) {
  // tanh approximation of GELU
  const float c = 0.7978845608028654f;   // sqrt(2/pi)
  const float u = c * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanh_fast(u));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float x2 = x * x;
  const float u = c * (x + 0.044715f * x2 * x);
  const float t = tanh_fast(u);
  const float du = c * (1.0f + 3.0f * 0.044715f * x2);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[TM * LDS_K];
  __shared__ __attribute__((aligned(16))) __bf16 sB[TN * LDS_K];   // stored [n][k]

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;   // wavefront quadrant

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += TK) {
    // stage A[m0:m0+64, k0:k0+32] and B[k0:k0+32, n0:n0+64]: 2048 elements each, 8 per thread
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;
      const int r = idx / TK, kk = idx % TK;                    // A: row r, column kk
      const int m = m0 + r, k = k0 + kk;
      __bf16 va = (__bf16)0.0f;
      if (m < g.M && k < g.K) va = g.A[(size_t)m * g.sAm + (size_t)k * g.sAk];
      sA[r * LDS_K + kk] = va;
      const int c = idx / TK, kb = idx % TK;                    // B: column c, row kb
      const int n = n0 + c, kB = k0 + kb;
      __bf16 vb = (__bf16)0.0f;
      if (n < g.N && kB < g.K) vb = g.B[(size_t)kB * g.sBk + (size_t)n * g.sBn];
      sB[c * LDS_K + kb] = vb;
    }
    __syncthreads();
    // v_mfma_f32_16x16x32_bf16: lane holds A row (lane % 16), k = 8 (lane / 16) .. +7, and
    // B column (lane % 16), the same k range.
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&sA[(wm + 16 * i + r16) * LDS_K + kq]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[(wn + 16 * j + r16) * LDS_K + kq]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // D layout of 16x16 MFMA: lane holds column (lane % 16), rows 4 (lane / 16) + r, r = 0..3
  const int cn = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + rq + r, n = n0 + wn + 16 * j + cn;
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][r];
        const size_t o = (size_t)m * g.sCm + (size_t)n * g.sCn;
        switch (g.epilogue) {
          case kEpiStore:
            g.C[o] = g.accumulate ? g.C[o] + v : v;
            break;
          case kEpiBiasGelu: {
            v += g.bias[n];
            g.C[o] = v;                                        // pre-activation (fp32)
            g.Cb[(size_t)m * g.sCbm + n] = (__bf16)gelu_f(v);  // activation (bf16)
            break;
          }
          case kEpiBias:
            g.C[o] = v + g.bias[n];
            break;
        }
      }
}

__global__ void gelu_backward_kernel(const float* __restrict__ pre, const float* __restrict__ dact,
                                     __bf16* __restrict__ dpre_bf, float* __restrict__ dpre, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = dact[i] * gelu_grad(pre[i]);
  if (dpre) dpre[i] = d;
  dpre_bf[i] = (__bf16)d;
}

__global__ void to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// column sums of a [M][N] fp32 matrix (bias gradients); grid over N, block reduces over M
__global__ void colsum_kernel(const float* __restrict__ x, int M, int N, float* __restrict__ out) {
  __shared__ float part[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += x[(size_t)m * N + n];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = part[0];
}

__global__ void adam_kernel(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1,
                            float b2, float eps, float bc1, float bc2, float wd) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

// Warm-start loss on the final layer output out[M][NOUT] = [coefficients (nsh) | logits (nlog)]:
//   L = mean_m [ sum_c (out_c - t_c)^2 * wc  +  wl * sum_r BCE(logit_r, label_r) ]
// writes dL/dout (fp32) and accumulates the two loss terms into loss[0..1].
__global__ void warm_loss_kernel(const float* __restrict__ out, const float* __restrict__ tgt_coef,
                                 const uint8_t* __restrict__ tgt_neg, int M, int nsh, int nlog, float wc,
                                 float wl, float* __restrict__ dout, float* __restrict__ loss) {
  const int m = blockIdx.x;
  const int NOUT = nsh + nlog;
  float lc = 0.f, ll = 0.f;
  for (int c = threadIdx.x; c < NOUT; c += blockDim.x) {
    const float o = out[(size_t)m * NOUT + c];
    float d;
    if (c < nsh) {
      const float e = o - tgt_coef[(size_t)m * nsh + c];
      lc += e * e;
      d = 2.f * wc * e;
    } else {
      const float y = (float)tgt_neg[(size_t)m * nlog + (c - nsh)];
      const float p = 1.f / (1.f + expf(-o));
      // BCE with logits, numerically stable form
      ll += fmaxf(o, 0.f) - o * y + log1pf(expf(-fabsf(o)));
      d = wl * (p - y);
    }
    dout[(size_t)m * NOUT + c] = d / (float)M;
  }
  __shared__ float s1[256], s2[256];
  s1[threadIdx.x] = lc;
  s2[threadIdx.x] = ll;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(loss + 0, wc * s1[0] / (float)M);
    atomicAdd(loss + 1, wl * s2[0] / (float)M);
  }
}

// x[m][i] = dw[m][dwis[i]] / s0[m] in bf16, s0 = mean of the b=0 volumes (1 if not positive).
__global__ void features_kernel(const float* __restrict__ dw, int ld, size_t M, const int* __restrict__ dwis,
                                int n_in, const int* __restrict__ b0, int n_b0, __bf16* __restrict__ x,
                                float* __restrict__ s0out) {
  const size_t m = (size_t)blockIdx.x * 4 + threadIdx.y;
  if (m >= M) return;
  const float* row = dw + m * (size_t)ld;
  float s = 0.f;
  for (int j = 0; j < n_b0; ++j) s += row[b0[j]];
  float s0 = n_b0 ? s / (float)n_b0 : 1.f;
  if (!(s0 > 0.f) || !isfinite(s0)) s0 = 1.f;
  const float inv = 1.f / s0;
  for (int i = threadIdx.x; i < n_in; i += blockDim.x) {
    float v = row[dwis[i]] * inv;
    if (!isfinite(v)) v = 0.f;
    x[m * (size_t)n_in + i] = (__bf16)v;
  }
  if (threadIdx.x == 0) s0out[m] = s0;
}

__global__ void propose_kernel(const float* __restrict__ out, const float* __restrict__ s0, size_t M, int nsh,
                               int nlog, float* __restrict__ F0, unsigned long long* __restrict__ negmask) {
  const size_t m = (size_t)blockIdx.x;
  if (m >= M) return;
  const int nout = nsh + nlog, words = (nlog + 63) / 64;
  const float* o = out + m * (size_t)nout;
  for (int c = threadIdx.x; c < nsh; c += blockDim.x) F0[m * (size_t)nsh + c] = o[c] * s0[m];
  for (int w = threadIdx.x; w < words; w += blockDim.x) {
    unsigned long long bits = 0ull;
    for (int b = 0; b < 64; ++b) {
      const int r = w * 64 + b;
      if (r < nlog && o[nsh + r] > 0.f) bits |= 1ull << b;
    }
    negmask[m * (size_t)words + w] = bits;
  }
}

// ---- fused inference: features -> 4 layers -> proposal, one 64-voxel tile per workgroup ----
// Activations stay in LDS (bf16, two ping-pong buffers); weights are read from L2 as the B operand
// of v_mfma_f32_16x16x32_bf16 from a transposed [fan_out][fan_in] copy (16-byte loads of eight
// consecutive k), prefetched one k-step ahead.  The arithmetic is the unfused path's: the same MFMA
// over k in steps of 32, bias, GELU and rounding to bf16, so both produce identical proposals.
constexpr int FI_TM = 64;               // voxels per tile
constexpr int FI_PH = kInferHidden + 8; // LDS row pitch (bf16) of hidden activations
constexpr int FI_PX = kInferMaxIn + 8;  // LDS row pitch (bf16) of network inputs

// acc = in[64 x K] * W[K x (n0 + 64 wave .. + 63)] for the column blocks that start below N.
__device__ __forceinline__ void fi_chunk(const __bf16* sIn, int pin, int K, const __bf16* __restrict__ WT,
                                         int ldw, int N, int n0, int wave, int lane, f32x4 (&acc)[4][4]) {
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int nb = n0 + 64 * wave;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero = {};
  bf16x8 b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    b[j] = (nb + 16 * j < N) ? *reinterpret_cast<const bf16x8*>(WT + (size_t)(nb + 16 * j + r16) * ldw + kq) : zero;
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 bn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      bn[j] = (k0 + 32 < K && nb + 16 * j < N)
                  ? *reinterpret_cast<const bf16x8*>(WT + (size_t)(nb + 16 * j + r16) * ldw + k0 + 32 + kq)
                  : zero;
    bf16x8 a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(sIn + (16 * i + r16) * pin + k0 + kq);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (nb + 16 * j < N)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = bn[j];
  }
}

// hidden layer epilogue: out = bf16(gelu(acc + bias)) into the LDS activation buffer
__device__ __forceinline__ void fi_store_hidden(const f32x4 (&acc)[4][4], const float* __restrict__ bias, __bf16* sOut,
                                                int n0, int wave, int lane) {
  const int cn = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + 64 * wave + 16 * j + cn;
    const float bn = bias[n];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) sOut[(16 * i + rq + r) * FI_PH + n] = (__bf16)gelu_f(acc[i][j][r] + bn);
  }
}

__global__ __launch_bounds__(256) void infer_propose_kernel(InferArgs a) {
  static_assert(FI_PX <= FI_PH, "the input tile lives in the second activation buffer");
  __shared__ __attribute__((aligned(16))) __bf16 sHa[FI_TM * FI_PH];
  __shared__ __attribute__((aligned(16))) __bf16 sHb[FI_TM * FI_PH];
  __shared__ float sS0[FI_TM];
  __bf16* const sX = sHb;   // network inputs; dead once the first layer has read them
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const size_t m0 = (size_t)blockIdx.x * FI_TM;

  // features: x = dw[dwis] / s0, s0 = mean of the b=0 volumes (1 when not positive)
  if (tid < FI_TM) {
    const size_t m = m0 + tid;
    float s0 = 1.f;
    if (m < a.M) {
      const float* row = a.dw + m * (size_t)a.ld;
      float s = 0.f;
      for (int j = 0; j < a.n_b0; ++j) s += row[a.b0[j]];
      s0 = a.n_b0 ? s / (float)a.n_b0 : 1.f;
      if (!(s0 > 0.f) || !isfinite(s0)) s0 = 1.f;
    }
    sS0[tid] = s0;
  }
  __syncthreads();
  for (int e = tid; e < FI_TM * a.k0; e += 256) {
    const int row = e / a.k0, i = e % a.k0;
    const size_t m = m0 + row;
    float v = 0.f;
    if (m < a.M && i < a.n_in) {
      v = a.dw[m * (size_t)a.ld + a.dwis[i]] * (1.f / sS0[row]);
      if (!isfinite(v)) v = 0.f;
    }
    sX[row * FI_PX + i] = (__bf16)v;
  }
  __syncthreads();

  f32x4 acc[4][4];
  fi_chunk(sX, FI_PX, a.k0, a.WT[0], a.ldw[0], kInferHidden, 0, wave, lane, acc);
  fi_store_hidden(acc, a.bias[0], sHa, 0, wave, lane);
  __syncthreads();
  fi_chunk(sHa, FI_PH, kInferHidden, a.WT[1], a.ldw[1], kInferHidden, 0, wave, lane, acc);
  fi_store_hidden(acc, a.bias[1], sHb, 0, wave, lane);
  __syncthreads();
  fi_chunk(sHb, FI_PH, kInferHidden, a.WT[2], a.ldw[2], kInferHidden, 0, wave, lane, acc);
  fi_store_hidden(acc, a.bias[2], sHa, 0, wave, lane);
  __syncthreads();

  // output layer: coefficients (n < nsh) -> F0 = out * s0; logits -> one flag byte per direction in
  // sHb (free now), packed into 64-bit words below
  unsigned char* flags = reinterpret_cast<unsigned char*>(sHb);
  const int nout = a.nsh + a.nlog;
  const int fpitch = (a.nlog + 63) / 64 * 64;
  const int cn = lane & 15, rq = (lane >> 4) * 4;
  for (int n0 = 0; n0 < nout; n0 += 4 * 64) {
    fi_chunk(sHa, FI_PH, kInferHidden, a.WT[3], a.ldw[3], nout, n0, wave, lane, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 64 * wave + 16 * j + cn;
      if (n >= nout) continue;
      const float bn = a.bias[3][n];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * i + rq + r;
          const float v = acc[i][j][r] + bn;
          if (n < a.nsh) {
            const size_t m = m0 + row;
            if (a.F0 && m < a.M) a.F0[m * (size_t)a.nsh + n] = v * sS0[row];
          } else {
            flags[row * fpitch + (n - a.nsh)] = v > 0.f ? 1 : 0;
          }
        }
    }
  }
  __syncthreads();
  const int words = (a.nlog + 63) / 64;
  for (int e = tid; e < FI_TM * words; e += 256) {
    const int row = e / words, w = e % words;
    const size_t m = m0 + row;
    if (m >= a.M) continue;
    unsigned long long bits = 0ull;
    for (int b = 0; b < 64; ++b) {
      const int r = 64 * w + b;
      if (r < a.nlog && flags[row * fpitch + r]) bits |= 1ull << b;
    }
    a.negmask[m * (size_t)words + w] = bits;
  }
}

// WT[n][k] = W[k][n] (bf16), zero padded to [npad][kpad]
__global__ void transpose_pad_kernel(const __bf16* __restrict__ W, int fin, int fout, __bf16* __restrict__ WT,
                                     int kpad, int npad) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)kpad * npad) return;
  const int n = (int)(i / kpad), k = (int)(i % kpad);
  WT[i] = (n < fout && k < fin) ? W[(size_t)k * fout + n] : (__bf16)0.0f;
}

}  // namespace

hipError_t transpose_pad(const __bf16* W, int fin, int fout, __bf16* WT, int kpad, int npad, hipStream_t stream) {
  const size_t n = (size_t)kpad * npad;
  if (!n) return hipSuccess;
  transpose_pad_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(W, fin, fout, WT, kpad, npad);
  return hipGetLastError();
}

hipError_t infer_propose(const InferArgs& a, hipStream_t stream) {
  if (!a.M) return hipSuccess;
  if (a.k0 % 32 || a.k0 > kInferMaxIn || a.nsh + a.nlog > 512 || a.nlog > 512) return hipErrorInvalidValue;
  infer_propose_kernel<<<(unsigned)((a.M + FI_TM - 1) / FI_TM), 256, 0, stream>>>(a);
  return hipGetLastError();
}

hipError_t features(const float* dw, int ld, size_t M, const int* dwis, int n_in, const int* b0, int n_b0,
                    __bf16* x, float* s0, hipStream_t stream) {
  if (!M) return hipSuccess;
  features_kernel<<<(unsigned)((M + 3) / 4), dim3(64, 4), 0, stream>>>(dw, ld, M, dwis, n_in, b0, n_b0, x, s0);
  return hipGetLastError();
}

hipError_t propose(const float* out, const float* s0, size_t M, int nsh, int nlog, float* F0,
                   unsigned long long* negmask, hipStream_t stream) {
  if (!M) return hipSuccess;
  propose_kernel<<<(unsigned)M, 64, 0, stream>>>(out, s0, M, nsh, nlog, F0, negmask);
  return hipGetLastError();
}

hipError_t gemm_bf16(const GemmArgs& g, hipStream_t stream) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  dim3 grid((unsigned)((g.N + TN - 1) / TN), (unsigned)((g.M + TM - 1) / TM));
  gemm_bf16_kernel<<<grid, 256, 0, stream>>>(g);
  return hipGetLastError();
}

hipError_t gelu_backward(const float* pre, const float* dact, __bf16* dpre_bf, float* dpre, size_t n,
                         hipStream_t stream) {
  if (!n) return hipSuccess;
  gelu_backward_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(pre, dact, dpre_bf, dpre, n);
  return hipGetLastError();
}

hipError_t to_bf16(const float* x, __bf16* y, size_t n, hipStream_t stream) {
  if (!n) return hipSuccess;
  to_bf16_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(x, y, n);
  return hipGetLastError();
}

hipError_t colsum(const float* x, int M, int N, float* out, hipStream_t stream) {
  if (N <= 0) return hipSuccess;
  colsum_kernel<<<N, 256, 0, stream>>>(x, M, N, out);
  return hipGetLastError();
}

hipError_t adam_step(float* w, __bf16* wb, const float* g, float* m, float* v, size_t n, float lr, float b1,
                     float b2, float eps, int t, float wd, hipStream_t stream) {
  if (!n) return hipSuccess;
  const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
  adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(w, wb, g, m, v, n, lr, b1, b2, eps, bc1, bc2, wd);
  return hipGetLastError();
}

hipError_t warm_loss(const float* out, const float* tgt_coef, const uint8_t* tgt_neg, int M, int nsh, int nlog,
                     float wc, float wl, float* dout, float* loss, hipStream_t stream) {
  if (M <= 0) return hipSuccess;
  warm_loss_kernel<<<M, 256, 0, stream>>>(out, tgt_coef, tgt_neg, M, nsh, nlog, wc, wl, dout, loss);
  return hipGetLastError();
}

}  // namespace mrcsd
