// Test infrastructure (tests/test_icls_kernel_emulated.py): a CPU emulation of one 64-lane
// wavefront per workgroup: each lane is a fiber (ucontext); every
// cross-lane operation and __syncthreads is a collective all 64 lanes must reach together (same
// kind and call site), otherwise the run stops and reports where the lanes are.
#pragma once
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost, hipMemcpyDefault };
typedef void* hipStream_t;
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
enum hipDeviceAttribute_t {
  hipDeviceAttributeMultiprocessorCount,
  hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
  hipDeviceAttributeMaxSharedMemoryPerBlock
};

namespace emu {
constexpr int W = 64;
struct Idx { unsigned x, y, z; };
extern int g_lane;
extern unsigned g_block, g_grid;
inline Idx tid() { return {unsigned(g_lane), 0u, 0u}; }
inline Idx bid() { return {g_block, 0u, 0u}; }
inline Idx gdim() { return {g_grid, 1u, 1u}; }
inline Idx bdim() { return {unsigned(W), 1u, 1u}; }
enum Kind { K_SYNC = 0, K_BALLOT, K_READLANE, K_READFIRST, K_SHFL };
// lane writes its value for the current collective, then waits for all lanes; returns the slot
// row of this collective (valid until this lane's next collective)
const uint64_t* collective(uint64_t v, int kind, const char* file, int line);
void run_workgroup(const std::function<void()>& body, unsigned block, unsigned grid);
extern long long g_collectives;
}  // namespace emu

#define threadIdx (emu::tid())
#define blockIdx (emu::bid())
#define gridDim (emu::gdim())
#define blockDim (emu::bdim())

#define __syncthreads() (void)emu::collective(0, emu::K_SYNC, __FILE__, __LINE__)
#define __ballot(p) emu_ballot((p), __FILE__, __LINE__)
#define __builtin_amdgcn_readlane(v, s) emu_readlane((v), (s), __FILE__, __LINE__)
#define __builtin_amdgcn_readfirstlane(v) emu_readfirst((v), __FILE__, __LINE__)
#define __shfl_xor(v, o) emu_shfl_xor((v), (o), __FILE__, __LINE__)
#define __shfl(v, s) emu_shfl((v), (s), __FILE__, __LINE__)
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt((double)(x)))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#define __ffsll(x) __builtin_ffsll(x)
#define __popcll(x) __builtin_popcountll(x)

inline unsigned long long emu_ballot(bool p, const char* f, int l) {
  const uint64_t* s = emu::collective(p ? 1 : 0, emu::K_BALLOT, f, l);
  unsigned long long m = 0;
  for (int i = 0; i < emu::W; ++i) m |= (s[i] & 1ull) << i;
  return m;
}
inline unsigned emu_readfirst(unsigned v, const char* f, int l) {
  const uint64_t* s = emu::collective(v, emu::K_READFIRST, f, l);
  return (unsigned)s[0];
}
inline int emu_readfirst(int v, const char* f, int l) { return (int)emu_readfirst((unsigned)v, f, l); }
// the lane index must be wave-uniform: taken from lane 0 as the hardware would (readfirstlane)
// and reported when it differs
void emu_report_nonuniform(const char* f, int l, int a, int b);
inline unsigned emu_readlane(unsigned v, int src, const char* f, int l) {
  const uint64_t* s = emu::collective(((uint64_t)(unsigned)src << 32) | v, emu::K_READLANE, f, l);
  const int s0 = (int)(s[0] >> 32);
  for (int i = 1; i < emu::W; ++i)
    if ((int)(s[i] >> 32) != s0) emu_report_nonuniform(f, l, s0, (int)(s[i] >> 32));
  if (s0 < 0 || s0 > 63) emu_report_nonuniform(f, l, s0, -1);
  return (unsigned)s[s0 & 63];
}
inline int emu_readlane(int v, int src, const char* f, int l) { return (int)emu_readlane((unsigned)v, src, f, l); }
template <class T>
inline T emu_shfl_xor(T v, int off, const char* f, int l) {
  static_assert(sizeof(T) <= 8, "");
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  const uint64_t* s = emu::collective(u, emu::K_SHFL, f, l);
  T r;
  const uint64_t w = s[(emu::g_lane ^ off) & 63];
  std::memcpy(&r, &w, sizeof(T));
  return r;
}
// each lane reads the value of lane `src` (per lane, any lane of the wavefront)
template <class T>
inline T emu_shfl(T v, int src, const char* f, int l) {
  static_assert(sizeof(T) <= 8, "");
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  const uint64_t* s = emu::collective(u, emu::K_SHFL, f, l);
  T r;
  const uint64_t w = s[src & 63];
  std::memcpy(&r, &w, sizeof(T));
  return r;
}
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long x) { double r; std::memcpy(&r, &x, 8); return r; }
template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p = o + v; return o; }
using std::max;
using std::min;
using std::isfinite;
using std::isnan;

// HIP runtime stand-ins over host memory
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::calloc(1, n ? n : 1); return *p ? hipSuccess : hipErrorOutOfMemory; }
template <class T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = 0) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = 0) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emu error"; }
// one CU; LDS as gfx950's 160 KB per CU and per workgroup
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  *v = a == hipDeviceAttributeMultiprocessorCount ? 1 : 163840;
  return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) { *free_b = *total_b = size_t(1) << 36; return hipSuccess; }
