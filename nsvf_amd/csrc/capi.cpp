// Library-wide C-ABI pieces: error text and version.
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace nsvf {

static thread_local char g_err[512] = {0};

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = 0; }

}  // namespace nsvf

extern "C" {

const char* nsvf_last_error(void) { return nsvf::g_err; }

int nsvf_version(void) { return 100; }

}  // extern "C"
