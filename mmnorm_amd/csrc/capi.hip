// Error plumbing and version query of the C ABI (include/mmnorm_amd.h).
#include <stdarg.h>
#include <stdio.h>

#include "mmn_internal.h"

static thread_local char g_last_error[512] = "";

void mmn_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int mmn_check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    mmn_set_error("%s: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

extern "C" int mmn_abi_version(void) { return MMN_ABI_VERSION; }

extern "C" const char *mmn_last_error(void) { return g_last_error; }
