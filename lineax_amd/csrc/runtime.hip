// Error reporting, version and device queries for the lineax_amd C ABI.
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "lxa_common.h"

namespace lxa {

namespace {
thread_local char g_last_error[512] = "";
}

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int hip_error(hipError_t e, const char* what) {
  return set_error(LXA_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int num_cus() {
  static std::mutex mu;
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

}  // namespace lxa

extern "C" const char* lxa_version(void) { return "lineax_amd 0.1.0 (gfx950)"; }

extern "C" const char* lxa_last_error(void) { return lxa::g_last_error; }
