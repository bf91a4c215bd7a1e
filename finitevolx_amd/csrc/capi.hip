// Library-wide C ABI: error reporting and version.
#include <cstdarg>
#include <cstring>

#include "fvx_common.hpp"

namespace fvx {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

}  // namespace fvx

extern "C" {

const char* fvx_last_error(void) { return fvx::last_error(); }

int fvx_abi_version(void) { return 1; }

int fvx_device_synchronize(void) {
  FVX_HIP(hipDeviceSynchronize());
  return 0;
}
}
