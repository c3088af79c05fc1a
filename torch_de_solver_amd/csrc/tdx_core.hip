// ABI version and the per-thread error message.
#include "tdx_common.h"

namespace tdx {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

}  // namespace tdx

extern "C" int tdx_version(void) { return 3; }
extern "C" const char *tdx_last_error(void) { return tdx::g_last_error.c_str(); }
