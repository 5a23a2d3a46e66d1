// internal.h - host-side helpers shared by the libpyrit translation units.
#pragma once
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "pyrit.h"
#include "pyrit_hip_build.h"

namespace pyr {

// Per-thread error message (the reference's diag.zig:17-35): cleared at every
// API entry, set by fail(), read by pyr_error_message().
extern thread_local char g_diag[512];

inline void diag_clear() { g_diag[0] = 0; }

inline PyrResult fail(PyrResult code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_diag, sizeof(g_diag), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace pyr
