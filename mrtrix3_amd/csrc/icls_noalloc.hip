// mricls_solve_noalloc: the batched ICLS solve (mricls_solve) for callers that cannot allocate
// device memory (code under stream capture, serving loops, allocation-free harnesses).  The solve's
// scratch -- the work counters of its passes and the two queues that carry voxels whose active set
// outgrows a pass to the next, larger one -- lives in a fixed device array of this file instead of
// buffers the context grows.  Each device holds its own copy; the host reaches it only through
// the address hipGetSymbolAddress returns for the current device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/mricls.h"
#include "capi_icls_ctx.h"

namespace {
constexpr size_t kCounterBytes = 256;   // 2 x kIclsMaxPasses counters, padded to the queue alignment
constexpr size_t kQueueBytes = size_t(MRICLS_NOALLOC_MAX_VOXELS) * sizeof(uint32_t);
static_assert(2 * mrcsd::kIclsMaxPasses * sizeof(unsigned long long) <= kCounterBytes, "counters fit");
}  // namespace

// The scratch of mricls_solve_noalloc on each device (8 MiB + 256 B).
static __device__ __align__(256) unsigned char scratch[kCounterBytes + 2 * kQueueBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_ptr[kMaxDevices] = {};   // device address of `scratch`, per device

// The current device's copy of `scratch` (looked up once per device).
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return mrcsd::fail(MRCSD_ERR_ARG, "mricls_solve_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mrcsd::fail(MRCSD_ERR_HIP, std::string("mricls_solve_noalloc: hipGetSymbolAddress: ") +
                                            hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return MRCSD_OK;
}

}  // namespace

extern "C" int mricls_solve_noalloc(mricls_ctx* ctx, const float* b, int ld_b, const uint8_t* valid, size_t nvox,
                                    float* x, uint16_t* iters, uint16_t* nactive, uint64_t* stats, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_solve_noalloc: null context");
    if (nvox == 0) return MRCSD_OK;
    if (!b || !x) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_solve_noalloc: null b or x");
    if (ld_b < ctx->P.nmeas) return mrcsd::fail(MRCSD_ERR_ARG, "mricls_solve_noalloc: ld_b smaller than nmeas");
    if (nvox > size_t(MRICLS_NOALLOC_MAX_VOXELS))
      return mrcsd::fail(MRCSD_ERR_ARG, "mricls_solve_noalloc: " + std::to_string(nvox) +
                                            " voxels; at most " + std::to_string(MRICLS_NOALLOC_MAX_VOXELS) +
                                            " per call (split the batch)");
    MRCSD_HIP(hipSetDevice(ctx->device));
    unsigned char* base = nullptr;
    const int rc = scratch_address(ctx->device, &base);
    if (rc) return rc;
    mrcsd::IclsLaunch L{};
    L.b = b; L.ld_b = ld_b; L.valid = valid; L.nvox = nvox; L.x = x; L.iters = iters;
    L.nactive = nactive; L.stats = reinterpret_cast<unsigned long long*>(stats);
    mrcsd::IclsScratch S{};
    S.counters = reinterpret_cast<unsigned long long*>(base);
    S.queue[0] = reinterpret_cast<uint32_t*>(base + kCounterBytes);
    S.queue[1] = reinterpret_cast<uint32_t*>(base + kCounterBytes + kQueueBytes);
    S.capacity = MRICLS_NOALLOC_MAX_VOXELS;
    // the counters are cleared by a kernel at the start of the launch sequence, on `stream`
    MRCSD_HIP(mrcsd::icls_launch(ctx->P, L, S, ctx->capacity, static_cast<hipStream_t>(stream)));
    return MRCSD_OK;
  });
}
