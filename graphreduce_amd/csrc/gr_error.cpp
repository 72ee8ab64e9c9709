// Per-thread last-error message for the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "gr_internal.h"

namespace gr {

static thread_local char g_last_error[512] = "";

int set_error(int code, const char* msg) {
  snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
  return code;
}

int set_errorf(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace gr

extern "C" const char* gr_last_error(void) { return gr::g_last_error; }

extern "C" const char* gr_version(void) { return "graphreduce_amd 0.1.0 (gfx950)"; }
