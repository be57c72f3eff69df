// Thread-local error text and version for the C ABI.
#include <cstdarg>
#include <cstdio>
#include "common.hpp"

namespace {
thread_local char g_err[512] = "no error";
}

void maomi_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

MAOMI_API const char* maomi_last_error(void) { return g_err; }
MAOMI_API const char* maomi_version(void) { return "maomi_amd 0.1.0 (gfx950)"; }
