// terrain.hip - workload generator library (libpyrit_terrain.so): the seeded
// terrain of terrain.h emitted on the host (for tests and CPU baselines) or on
// the device (bench inputs and world chunks). Both emit the same voxels in the
// same order: chunk by chunk, columns x-fastest then z, cells bottom-up.
//
// Output layout ("packed" chunk batch): chunk c occupies
// voxels[offsets[c] .. offsets[c+1]) as PyrVoxel {x, y, z, attribute} in
// chunk-local coordinates; offsets is the exclusive prefix of the counts.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include "pyrit.h"
#include "terrain.h"

namespace {

struct Vox { int32_t x, y, z; uint32_t a; };

struct Gen {
    const PyrTerrainChunk* chunks;
    const PyrTerrainCave* caves;
    const uint32_t* bin_start;
    const uint32_t* bin_caves;
    const uint64_t* col_offsets; // exclusive prefix of columns per chunk, [count + 1]
    uint32_t count;
};

__device__ __forceinline__ uint32_t chunk_of(const Gen& g, uint64_t col) {
    uint32_t lo = 0, hi = g.count;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (g.col_offsets[mid] <= col) lo = mid; else hi = mid;
    }
    return lo;
}

struct NoEmit {
    __host__ __device__ void operator()(uint32_t, uint32_t, uint32_t, uint32_t) const {}
};

struct StoreEmit {
    Vox* out;
    uint64_t pos;
    __host__ __device__ void operator()(uint32_t x, uint32_t y, uint32_t z, uint32_t a) {
        out[pos++] = Vox{(int32_t)x, (int32_t)y, (int32_t)z, a};
    }
};

__global__ void __launch_bounds__(256) k_count(Gen g, uint64_t total_cols, uint32_t* col_counts) {
    uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= total_cols) return;
    uint32_t c = chunk_of(g, col);
    const PyrTerrainChunk* ch = &g.chunks[c];
    uint64_t local = col - g.col_offsets[c];
    uint32_t n = 1u << ch->log2_size;
    col_counts[col] = pyrgen_column(ch, g.caves, g.bin_start, g.bin_caves, (uint32_t)(local % n),
                                    (uint32_t)(local / n), NoEmit{});
}

__global__ void __launch_bounds__(256) k_emit(Gen g, uint64_t total_cols, const uint64_t* col_start, Vox* out) {
    uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= total_cols) return;
    uint32_t c = chunk_of(g, col);
    const PyrTerrainChunk* ch = &g.chunks[c];
    uint64_t local = col - g.col_offsets[c];
    uint32_t n = 1u << ch->log2_size;
    StoreEmit e{out, col_start[col]};
    pyrgen_column(ch, g.caves, g.bin_start, g.bin_caves, (uint32_t)(local % n), (uint32_t)(local / n), e);
}

// single-block exclusive scan of 64-bit counts, chunked (sizes here are at
// most a few million columns)
__global__ void __launch_bounds__(1024) k_scan(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* total) {
    __shared__ uint64_t part[1024];
    uint64_t per = (n + 1023) / 1024;
    uint64_t b = threadIdx.x * per, e = b + per < n ? b + per : n;
    uint64_t s = 0;
    for (uint64_t i = b; i < e; i++) s += in[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int i = 0; i < 1024; i++) {
            uint64_t v = part[i];
            part[i] = acc;
            acc += v;
        }
        *total = acc;
    }
    __syncthreads();
    uint64_t acc = part[threadIdx.x];
    for (uint64_t i = b; i < e; i++) {
        out[i] = acc;
        acc += in[i];
    }
}

__global__ void k_chunk_offsets(const uint64_t* col_start, const uint64_t* col_offsets, uint32_t count,
                                const uint64_t* total, uint64_t* offsets) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < count) offsets[c] = col_start[col_offsets[c]];
    if (c == count) offsets[count] = *total;
}

// world chunks: one thread per (chunk, column), voxels appended with an
// atomic counter per chunk into the generator's capacity layout
struct CapEmit {
    PyrVoxel* out;
    uint32_t* count;
    uint32_t capacity;
    __device__ void operator()(uint32_t x, uint32_t y, uint32_t z, uint32_t a) const {
        const uint32_t k = atomicAdd(count, 1u);
        if (k < capacity) out[k] = PyrVoxel{(int32_t)x, (int32_t)y, (int32_t)z, a};
    }
};

__global__ void __launch_bounds__(256) k_world(PyrTerrainWorld w, const PyrChunkKey* keys, uint32_t count,
                                               uint32_t chunk_log2, uint32_t capacity, PyrVoxel* voxels,
                                               uint32_t* counts) {
    const uint64_t cols = (uint64_t)1 << (2 * chunk_log2);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols * count) return;
    const uint32_t c = (uint32_t)(i / cols), col = (uint32_t)(i % cols);
    const PyrChunkKey k = keys[c];
    const uint32_t n = 1u << chunk_log2;
    CapEmit e{voxels + (uint64_t)c * capacity, &counts[c], capacity};
    pyrgen_world_column(&w, k.x, k.y, k.z, k.lod, chunk_log2, col % n, col / n, e);
}

struct HostEmit {
    PyrVoxel* out;
    uint32_t capacity;
    uint32_t* k;
    void operator()(uint32_t x, uint32_t y, uint32_t z, uint32_t a) const {
        if (*k < capacity && out) out[*k] = PyrVoxel{(int32_t)x, (int32_t)y, (int32_t)z, a};
        (*k)++;
    }
};

} // namespace

extern "C" {

// PyrWorldGenFn of the streamed-world API for the terrain above; `user`
// points to a PyrTerrainWorld in host memory.
__attribute__((visibility("default"))) void pyrgen_world_generate(void* user, const PyrWorldGenParams* p,
                                                                  void* stream) {
    const PyrTerrainWorld w = *(const PyrTerrainWorld*)user;
    const uint64_t threads = ((uint64_t)1 << (2 * p->chunk_log2)) * p->count;
    if (threads == 0) return;
    hipLaunchKernelGGL(k_world, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream, w,
                       (const PyrChunkKey*)(uintptr_t)p->chunks, p->count, p->chunk_log2, p->capacity,
                       (PyrVoxel*)(uintptr_t)p->voxels, (uint32_t*)(uintptr_t)p->counts);
}

// The same chunk on the host (columns z-major, x fastest, cells bottom-up);
// writes at most `capacity` voxels and returns how many the chunk has.
__attribute__((visibility("default"))) uint32_t pyrgen_world_chunk_host(const PyrTerrainWorld* w,
                                                                        const PyrChunkKey* key,
                                                                        uint32_t chunk_log2, PyrVoxel* out,
                                                                        uint32_t capacity) {
    const uint32_t n = 1u << chunk_log2;
    uint32_t k = 0;
    HostEmit e{out, capacity, &k};
    for (uint32_t lz = 0; lz < n; lz++)
        for (uint32_t lx = 0; lx < n; lx++)
            pyrgen_world_column(w, key->x, key->y, key->z, key->lod, chunk_log2, lx, lz, e);
    return k;
}

// Host generation. col_offsets[count+1] = exclusive prefix of 4^log2 columns.
// voxels == NULL: only offsets[count+1] (voxel prefix per chunk) is filled.
__attribute__((visibility("default"))) int pyrgen_generate_host(const PyrTerrainChunk* chunks, uint32_t count,
                                                                const PyrTerrainCave* caves,
                                                                const uint32_t* bin_start,
                                                                const uint32_t* bin_caves, uint64_t* offsets,
                                                                void* voxels) {
    uint64_t pos = 0;
    for (uint32_t c = 0; c < count; c++) {
        offsets[c] = pos;
        const PyrTerrainChunk* ch = &chunks[c];
        uint32_t n = 1u << ch->log2_size;
        for (uint32_t lz = 0; lz < n; lz++)
            for (uint32_t lx = 0; lx < n; lx++) {
                if (voxels) {
                    StoreEmit e{(Vox*)voxels, pos};
                    pos += pyrgen_column(ch, caves, bin_start, bin_caves, lx, lz, e);
                } else {
                    pos += pyrgen_column(ch, caves, bin_start, bin_caves, lx, lz, NoEmit{});
                }
            }
    }
    offsets[count] = pos;
    return 0;
}

// Surface height of a chunk's terrain at world column (x, z) (solid: y < h).
__attribute__((visibility("default"))) int32_t pyrgen_height_host(const PyrTerrainChunk* chunk, int32_t x, int32_t z) {
    return pyrgen_height(chunk, x, z);
}

// Device generation, two passes (count, emit). All pointers are device
// pointers; scratch holds 12 bytes per column (see pyrgen_scratch_bytes).
// voxels == NULL: counts only (offsets filled). Asynchronous on `stream`.
__attribute__((visibility("default"))) uint64_t pyrgen_scratch_bytes(uint64_t total_columns) {
    return total_columns * 12 + 256;
}

__attribute__((visibility("default"))) int pyrgen_generate_device(const PyrTerrainChunk* chunks, uint32_t count,
                                                                  const PyrTerrainCave* caves,
                                                                  const uint32_t* bin_start,
                                                                  const uint32_t* bin_caves,
                                                                  const uint64_t* col_offsets,
                                                                  uint64_t total_columns, void* scratch,
                                                                  uint64_t* offsets, void* voxels, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (count == 0 || total_columns == 0) return -1;
    Gen g{chunks, caves, bin_start, bin_caves, col_offsets, count};
    uint32_t* col_counts = (uint32_t*)scratch;
    uint64_t* col_start = (uint64_t*)((char*)scratch + ((total_columns * 4 + 255) / 256) * 256);
    uint64_t* total = col_start + total_columns;
    uint64_t blocks = (total_columns + 255) / 256;
    hipLaunchKernelGGL(k_count, dim3((uint32_t)blocks), dim3(256), 0, s, g, total_columns, col_counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, col_counts, total_columns, col_start, total);
    hipLaunchKernelGGL(k_chunk_offsets, dim3((count + 256) / 256), dim3(256), 0, s, col_start, col_offsets,
                       count, total, offsets);
    if (voxels)
        hipLaunchKernelGGL(k_emit, dim3((uint32_t)blocks), dim3(256), 0, s, g, total_columns, col_start,
                           (Vox*)voxels);
    return hipGetLastError() == hipSuccess ? 0 : -6;
}

} // extern "C"
