// internal.h - host-side helpers shared by the libpyrit translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "pyrit.h"
#include "pyrit_hip_build.h"

// A host DAG (pyr_dag_*): the DAG arrays in host memory.
struct PyrDag {
    uint32_t log2_size = 0, root = 0;
    uint64_t voxel_count = 0;
    std::vector<uint32_t> nodes;
    std::vector<uint64_t> leaves;
    std::vector<uint32_t> attrs;
    bool has_attrs = true;
};

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

// Device memory of builds, build outputs and resident server generations
// comes from the library's stream-ordered cache (devcache.cpp): freed blocks
// stay mapped and serve later requests of their size class, reuse on another
// stream waits for the freeing stream on the device.
hipError_t pool_alloc(void** p, uint64_t bytes, hipStream_t s);
void pool_free(void* p, hipStream_t s);

// Build tuning of one workload: the caller's non-zero fields over the
// tuning installed for the class (log2_size, bit length of the record count,
// bit length of the chunk count) over the built-in defaults (tune.cpp).
PyrhipBuildTuning resolve_tuning(const PyrhipBuildTuning* caller, uint32_t log2_size, uint32_t n,
                                 uint32_t chunk_count);

// Context internals used by streamed worlds (api.cpp): a chunk batch is
// installed into the geometry pools once and every chunk gets a geometry
// that refers to the batch's shared ranges.
hipStream_t ctx_stream(PyrContext* ctx);
PyrResult ctx_install_batch(PyrContext* ctx, PyrhipBuilt* built, int32_t* batch_id);
void ctx_release_batch(PyrContext* ctx, int32_t batch_id);
uint64_t ctx_batch_bytes(PyrContext* ctx, int32_t batch_id);
PyrResult ctx_geometry_from_batch(PyrContext* ctx, int32_t batch_id, uint32_t root, uint32_t first_attr,
                                  uint32_t voxel_count, uint32_t log2_size, PyrGeometry* out);

}  // namespace pyr
