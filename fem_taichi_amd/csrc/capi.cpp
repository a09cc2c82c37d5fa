// Error reporting and device queries of the fem355 C ABI.
#include "common.hpp"

namespace fem {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace fem

extern "C" const char* fem_last_error(void) { return fem::g_err; }

extern "C" int fem_abi_version(void) { return 2; }

extern "C" int fem_device_count(int* count) {
  FEM_REQUIRE(count != nullptr, FEM_E_INVALID, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return FEM_OK;
}
