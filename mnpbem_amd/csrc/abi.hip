// C ABI plumbing: library version and the last-error string.
#include "common.h"

#include <mutex>
#include <string>

namespace {
thread_local std::string g_last_error;
}

void mnp_set_error(const std::string& msg) { g_last_error = msg; }

extern "C" int mnpbem_abi_version(void) { return 1; }

extern "C" const char* mnpbem_last_error(void) { return g_last_error.c_str(); }

extern "C" int mnpbem_device_sync(void) {
  MNP_HIP_CHECK(hipDeviceSynchronize());
  return 0;
}
