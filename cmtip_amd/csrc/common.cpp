// Error reporting shared by every entry point of the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace cmtip {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace cmtip

extern "C" const char* cmtip_last_error(void) { return cmtip::g_err; }

extern "C" int cmtip_version(void) { return 100; }
