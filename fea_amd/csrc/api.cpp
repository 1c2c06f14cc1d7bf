// Error reporting and version of the fea_amd C ABI.
#include <cstdarg>
#include <cstdio>

#include "../../include/fea_amd.h"

namespace fea {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace fea

extern "C" const char *fea_last_error(void) { return fea::g_err; }

extern "C" int fea_version(void) { return 1; }
