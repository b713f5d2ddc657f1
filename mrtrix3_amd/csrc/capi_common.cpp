#include "capi_common.h"

#include <string>

namespace mrcsd {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int require_gpu() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return fail(MRCSD_ERR_NOGPU, std::string("no HIP device available") +
                                     (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : ""));
  return MRCSD_OK;
}

}  // namespace mrcsd

extern "C" {

const char* mrcsd_last_error(void) { return mrcsd::g_last_error.c_str(); }

const char* mrcsd_version(void) { return "mrcsd 0.1 (gfx950)"; }

}  // extern "C"
