// Small wavefront-level helpers shared by the gfx950 kernels (wave size 64).
#pragma once

#include <hip/hip_runtime.h>

namespace mrcsd {

// A value the whole wavefront agrees on, moved to scalar registers so that loops and branches
// over it are scalar.
__device__ __forceinline__ unsigned long long uniform64(unsigned long long x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(x & 0xffffffffull));
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ int uniform32(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Broadcast lane `src`'s value to the wavefront (src must be wave-uniform).
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffull), src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int lane_bcast(int v, int src) { return __builtin_amdgcn_readlane(v, src); }

// Lane `src`'s 64-bit word as a wave-uniform value (src must be wave-uniform).
__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int src) {
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)(v & 0xffffffffull), src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

// Sum over the wavefront.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Single-wavefront workgroups: __syncthreads() keeps program order and drains the LDS queue;
// the barrier itself is elided by the compiler for one-wave groups.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// Lexicographic (value, index) minimum over the wavefront; every lane gets the result.
// Lower index wins ties, matching a first-occurrence scan in ascending index order.
__device__ __forceinline__ void wave_argmin(double& v, int& i) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(i, off);
    const bool take = (ov < v) || (ov == v && oi < i);
    v = take ? ov : v;
    i = take ? oi : i;
  }
}

}  // namespace mrcsd
