// Error state of the fqehip C ABI: one message per calling host thread.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "fqehip_common.h"

namespace fqehip {

static thread_local char g_error[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

void clear_error() { g_error[0] = '\0'; }

}  // namespace fqehip

extern "C" const char* fqehip_last_error(void) { return fqehip::g_error; }

extern "C" int fqehip_version(void) { return 1; }
