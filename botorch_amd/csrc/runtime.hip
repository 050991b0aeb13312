// Library-level helpers: error text, device properties, version.
#include "bta_common.h"

#include <cstdarg>
#include <cstdio>
#include <mutex>

namespace bta {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return BTA_ERR_HIP;
}

// Compute units of the current device (cached per device id).
int num_cus() {
  static std::mutex& mu = *new std::mutex;  // never destroyed
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (cache[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cache[dev] = v;
  }
  return cache[dev];
}

// A 4 KiB pinned host buffer per calling thread, allocated on first use and
// never freed (freeing it from a thread-exit destructor could run after the
// HIP runtime has shut down), for the small copies of calls that synchronise
// before they return; nullptr if the allocation fails (callers then copy
// through pageable memory).
double* pinned_host_buffer() {
  static thread_local double* p = nullptr;
  static thread_local bool tried = false;
  if (!tried) {
    tried = true;
    void* v = nullptr;
    if (hipHostMalloc(&v, 4096, hipHostMallocDefault) == hipSuccess) p = (double*)v;
  }
  return p;
}

}  // namespace bta

extern "C" const char* bta_version(void) { return "botorch_amd 0.1.0 (gfx950)"; }
extern "C" int bta_abi_version(void) { return BTA_ABI_VERSION; }
extern "C" const char* bta_last_error(void) { return bta::g_err; }
