// Thread-local error text behind dfm_last_error().
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <cstdio>

namespace {
thread_local char g_err[512] = "";
}

namespace dfm {
int set_error(int code, const char* what, const char* detail) {
  if (detail)
    std::snprintf(g_err, sizeof(g_err), "dfm error %d: %s (%s)", code, what, detail);
  else
    std::snprintf(g_err, sizeof(g_err), "dfm error %d: %s", code, what);
  return code;
}
}  // namespace dfm

extern "C" const char* dfm_last_error(void) { return g_err; }
