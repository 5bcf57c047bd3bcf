// Status / error-string plumbing of the C ABI.
#include <stdarg.h>
#include <string.h>

#include "dem_common.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace dem {
int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}
}  // namespace dem

extern "C" {

const char* dem_last_error(void) { return g_last_error; }

int dem_version(void) { return 1; }

int dem_device_count(void) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return dem::set_error(DEM_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return c;
}

}  // extern "C"
