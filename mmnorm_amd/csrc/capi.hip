// Error plumbing and version query of the C ABI (include/mmnorm_amd.h), and
// the zero fill the entry points use instead of hipMemsetAsync (a kernel
// node, so that a caller may capture the entry points in a HIP graph).
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>

#include "mmn_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_zero_words(uint32_t *__restrict__ p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

}  // namespace

int mmn_zero_async(void *p, size_t bytes, hipStream_t st, const char *what) {
  MMN_REQUIRE(bytes % 4 == 0 && ((uintptr_t)p & 3) == 0, MMN_ERR_ARG,
              "%s: zero fill of %zu bytes is not whole aligned 32-bit words", what, bytes);
  if (bytes == 0) return 0;
  const size_t n = bytes / 4;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, st, (uint32_t *)p, n);
  return mmn_check_launch(what);
}

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
