// Error state and small library-wide entry points.
#include "common.h"

#include <string.h>

namespace mk {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  return MK_OK;
}

}  // namespace mk

extern "C" const char* mk_last_error(void) { return mk::g_err; }
extern "C" int mk_version(void) { return 1; }
