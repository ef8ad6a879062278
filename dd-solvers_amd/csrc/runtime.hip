// Library-wide entry points: ABI version and per-thread error messages.
#include <cstdarg>

#include "common.hpp"

namespace dds {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

}  // namespace dds

extern "C" int dds_abi_version(void) { return DDS_ABI_VERSION; }

extern "C" const char* dds_last_error(void) { return dds::g_error; }
