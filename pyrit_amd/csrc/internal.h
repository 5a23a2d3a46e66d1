// internal.h - host-side helpers shared by the libpyrit translation units.
#pragma once
#include <hip/hip_runtime.h>
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

// Device memory of builds and geometries comes from the device's stream-ordered
// pool, configured once to keep freed memory (release threshold = max): a build
// allocates tens of GB of outputs and scratch, and recycling them from the
// pool replaces page mapping and unmapping on every call.
bool pool_ready(int device);
hipError_t pool_alloc(void** p, uint64_t bytes, hipStream_t s);
void pool_free(void* p, hipStream_t s);

}  // namespace pyr
