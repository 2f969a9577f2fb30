// Error reporting for the C ABI: every entry point returns an int status (0 = ok) and leaves
// a message for mgpbd_last_error() in thread-local storage; no exception crosses the ABI.
#include <string>

#include "mgpbd_internal.h"

static thread_local std::string g_last_error;

void mgpbd_set_error(const std::string& msg) { g_last_error = msg; }

int mgpbd_fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

extern "C" const char* mgpbd_last_error(void) { return g_last_error.c_str(); }

extern "C" int mgpbd_abi_version(void) { return MGPBD_ABI_VERSION; }
