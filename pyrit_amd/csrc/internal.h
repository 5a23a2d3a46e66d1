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
// API entry, set by fail(), read by pyr_error_message(). The entry also clears
// the thread's HIP error state: launches are checked with hipGetLastError,
// which would otherwise report an earlier failed call of the thread (one
// whose result was already handled) as the launch's own error.
extern thread_local char g_diag[512];

inline void diag_clear() {
    g_diag[0] = 0;
    (void)hipGetLastError();
}

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
// whether such a build takes the bucket finish under the resolved tuning (build.hip)
bool bucket_finish(const PyrhipBuildTuning& tun, uint32_t log2_size, uint64_t n, uint32_t chunk_count);

// Palette packing of a chunk batch's attributes (palette.hip). lo holds the
// first attribute rank of each chunk and, last, the total (k + 1 entries).
// palette_scan reads back per chunk the number of distinct attributes or
// kPaletteOverflow (more than 16: 32-bit storage); palette_words gives a
// chunk's words in the attribute pool; palette_pack writes chunk j at word
// dst[j] of pool: 16 palette entries then 4-bit indices, eight per word, or
// the 32-bit attributes. The scratch lives from scan to release.
constexpr uint32_t kPaletteOverflow = 0xFFFFFFFFu;
struct PaletteScratch {
    uint32_t* lo = nullptr;
    uint32_t* palettes = nullptr;
    uint32_t* table = nullptr;
    uint32_t* counts = nullptr;  // inside table's block
};
uint64_t palette_words(uint32_t count, uint64_t voxels);
PyrResult palette_scan(const uint32_t* attrs, const std::vector<uint32_t>& lo, hipStream_t s,
                       std::vector<uint32_t>* counts, PaletteScratch* scratch);
PyrResult palette_pack(const uint32_t* attrs, const std::vector<uint32_t>& lo, const std::vector<uint64_t>& dst,
                       uint32_t* pool, hipStream_t s, PaletteScratch* scratch);
void palette_release(PaletteScratch* scratch, hipStream_t s);

// Path edits (pathedit.hip): voxel edits applied to a finished DAG by
// writing new copies of the touched bricks and nodes into arena space. A
// job's edits are edits[edit_first ..][4] (x, y, z, attribute; chunk-local),
// sorted by brick Morton key with their order kept inside a brick
// (sort_path_edits). Node and leaf references inside the DAG are relative
// to node_base / leaf_base; arenas, attribute offsets and palette_src are
// offsets in the pools passed to run_path_edits. status: 0 applied
// (new_root, new_count valid), kPathEmpty, kPathOverflow (nothing usable).
constexpr uint32_t kPathEmpty = 2;     // the chunk became empty
constexpr uint32_t kPathOverflow = 4;  // the arena space given to the job was too small
struct PathEditJob {
    uint32_t root = 0, log2_size = 0;
    uint32_t node_base = 0, leaf_base = 0;
    uint32_t node_arena = 0, node_end = 0;  // arena words [node_arena, node_end)
    uint32_t leaf_arena = 0, leaf_end = 0;
    uint32_t attr_src = 0, attr_count = 0;
    uint32_t palette_src = 0xFFFFFFFFu;  // else: source attributes are 4-bit indices into 16 entries here
    uint32_t attr_dst = 0;               // new_count words
    uint32_t edit_first = 0, edit_count = 0;
    uint32_t new_root = 0, new_count = 0, status = 0, bricks = 0;
};
void sort_path_edits(const uint32_t* edits, uint32_t count, std::vector<uint32_t>* out);
// arena node words and leaves a job of sorted edits can need at most
void path_edit_capacity(const uint32_t* sorted, uint32_t count, uint32_t log2_size, uint64_t* node_words,
                        uint64_t* leaves);
PyrResult run_path_edits(std::vector<PathEditJob>& jobs, const std::vector<uint32_t>& edits, uint32_t* nodes_pool,
                         uint64_t* leaves_pool, const uint32_t* src_pool, uint32_t* dst_pool, hipStream_t s);

// Context internals used by streamed worlds (api.cpp): a chunk batch is
// installed into the geometry pools once and every chunk gets a geometry
// that refers to the batch's shared ranges.
hipStream_t ctx_stream(PyrContext* ctx);
// Chunk j of the k non-empty chunks of the batch owns attribute ranks
// [first[j], first[j + 1]) (first[k] = voxel count); on success attr_word[j]
// is its attribute offset relative to the batch's attribute range and
// palette[j] says whether it is palette packed.
PyrResult ctx_install_batch(PyrContext* ctx, PyrhipBuilt* built, const std::vector<uint32_t>& first,
                            int32_t* batch_id, std::vector<uint64_t>* attr_word, std::vector<uint8_t>* palette);
void ctx_release_batch(PyrContext* ctx, int32_t batch_id);
uint64_t ctx_batch_bytes(PyrContext* ctx, int32_t batch_id);
PyrResult ctx_geometry_from_batch(PyrContext* ctx, int32_t batch_id, uint32_t root, uint64_t attr_word,
                                  bool palette, uint32_t voxel_count, uint32_t log2_size, PyrGeometry* out);

// Path edits of batch chunks (context.zig:1093+): request i applies
// edits[first .. first + count)[4] (x, y, z, attribute; chunk-local, in the
// order given) to its geometry. applied[i] = 1 and voxels[i] = the new voxel
// count on success; a request that does not fit (arena, attribute pool, too
// many edits) or that empties its chunk is not applied and its chunk must be
// rebuilt. The arena restarts once no geometry uses it.
struct PathEditRequest {
    PyrGeometry geometry;
    uint32_t first, count;
};
constexpr uint32_t kPathEditMax = 4096;  // more edits in one chunk: a rebuild is cheaper (types.zig:602)
PyrResult ctx_path_edit(PyrContext* ctx, const std::vector<PathEditRequest>& reqs, const std::vector<uint32_t>& edits,
                        std::vector<uint8_t>* applied, std::vector<uint32_t>* voxels);
// three quarters of the node or leaf arena used (context.zig:1063-1068)
bool ctx_arena_pressure(PyrContext* ctx);

}  // namespace pyr
