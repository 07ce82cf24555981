// Library-wide C-ABI entry points: error reporting, version, device helpers.
#include "nls.h"
#include "nls_common.hpp"

#include <cstring>
#include <string>

namespace nls {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
} // namespace nls

extern "C" {

const char *nls_last_error(void) { return nls::g_last_error.c_str(); }

int nls_version(void) { return 100; }

int nls_device_info(char *name, int name_len, int *cu_count, size_t *hbm_bytes) {
  return nls::guard([&] {
    int dev = 0;
    NLS_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NLS_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
      std::strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
      name[name_len - 1] = '\0';
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  });
}

int nls_dev_malloc(void **ptr, size_t bytes) {
  return nls::guard([&] {
    NLS_REQUIRE(ptr, "nls_dev_malloc: null out pointer");
    NLS_HIP(hipMalloc(ptr, bytes));
  });
}
int nls_dev_free(void *ptr) { return nls::guard([&] { NLS_HIP(hipFree(ptr)); }); }
int nls_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return nls::guard([&] { NLS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int nls_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return nls::guard([&] { NLS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int nls_dev_synchronize(void) { return nls::guard([&] { NLS_HIP(hipDeviceSynchronize()); }); }

} // extern "C"
