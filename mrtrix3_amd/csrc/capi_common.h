// Shared plumbing of the C ABI: error strings, exception fencing, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/mrcsd.h"

namespace mrcsd {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

// A failed HIP call inside an ABI function: thrown and turned into MRCSD_ERR_HIP by guarded().
struct HipFailure : std::runtime_error {
  explicit HipFailure(const std::string& m) : std::runtime_error(m) {}
};

#define MRCSD_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::mrcsd::HipFailure(std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

// Runs an ABI body, translating exceptions into status codes.
template <class F>
int guarded(F&& body) {
  try {
    return body();
  } catch (const HipFailure& e) {
    return fail(MRCSD_ERR_HIP, e.what());
  } catch (const std::bad_alloc&) {
    return fail(MRCSD_ERR_SETUP, "out of host memory");
  } catch (const std::exception& e) {
    return fail(MRCSD_ERR_SETUP, e.what());
  } catch (...) {
    return fail(MRCSD_ERR_SETUP, "unknown error");
  }
}

// Fails with MRCSD_ERR_NOGPU unless a HIP device is present.
int require_gpu();

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { reset(); }

  void alloc(size_t n) {
    reset();
    if (n == 0) return;
    void* p = nullptr;
    MRCSD_HIP(hipMalloc(&p, n * sizeof(T)));
    ptr_ = static_cast<T*>(p);
    n_ = n;
  }
  void upload(const T* src, size_t n) {
    alloc(n);
    if (n) MRCSD_HIP(hipMemcpy(ptr_, src, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void download(T* dst, size_t n) const {
    if (n) MRCSD_HIP(hipMemcpy(dst, ptr_, n * sizeof(T), hipMemcpyDeviceToHost));
  }
  void reset() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  T* get() const { return ptr_; }
  size_t size() const { return n_; }

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
};

}  // namespace mrcsd
