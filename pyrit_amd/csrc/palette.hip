// palette.hip - compact attributes of a chunk batch (the reference's
// src/device/palette.zig and context.zig:1311-1370): a chunk whose voxels
// carry at most 16 distinct attributes stores a 16-entry palette followed by
// 4-bit indices, eight to a word (index of rank r in bits 4 (r & 7) of word
// r >> 3); other chunks keep 32 bits per voxel.
//
// MI355X shape: the reference runs one block per chunk for both steps, so the
// largest chunk (2^24 voxels at edge 256) sets the time. Here both steps run
// over tiles of a fixed number of attributes that never cross a chunk, found
// by a binary search of the block index in a per-chunk tile prefix:
//   k_palette_scan   per tile: distinct values in an LDS table (a register
//                    holds the last value seen, runs of one material cost no
//                    LDS traffic); at most 16 of them are merged into a
//                    32-slot table per chunk with global CAS; a chunk that
//                    passes 16 values is marked and its later tiles return
//   k_palette_final  per chunk: palette = the distinct values in ascending
//                    order (deterministic, unlike a hash-slot order), padded
//                    with its first entry; count or "overflow"
//   k_palette_pack   per tile of 2048 attributes: 256 index words (or the
//                    32-bit copy for overflow chunks); the chunk's first
//                    tile writes the palette in front
// Attribute 0xFFFFFFFF (a valid non-empty attribute) sends its chunk to
// 32-bit storage, since the tables use 0 as their empty slot and the value
// must survive exactly.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

constexpr int kPalBlock = 256;
constexpr uint32_t kPalEntries = 16;
constexpr uint32_t kScanPer = 16;                        // attributes per thread in a scan tile
constexpr uint32_t kScanTile = kPalBlock * kScanPer;     // 4096
constexpr uint32_t kPackTile = kPalBlock * 8;            // 2048 attributes = 256 index words
constexpr uint32_t kChunkSlots = 32;                     // global table per chunk
constexpr uint32_t kTileSlots = 64;                      // LDS table per tile

__device__ inline uint32_t slot_hash(uint32_t v) {
    uint32_t h = v * 0x9E3779B1u;
    return h ^ (h >> 15);
}

// chunk of block b: the last j with start[j] <= b (start has k + 1 entries)
__device__ inline uint32_t chunk_of(const uint32_t* start, uint32_t k, uint32_t b) {
    uint32_t lo = 0, hi = k;  // start[lo] <= b < start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kPalBlock) k_palette_scan(const uint32_t* attrs, const uint32_t* lo,
                                                             const uint32_t* tile_start, uint32_t k,
                                                             uint32_t* table, uint32_t* count, uint32_t* overflow) {
    __shared__ uint32_t s_keys[kTileSlots];
    __shared__ uint32_t s_count, s_full;
    const uint32_t c = chunk_of(tile_start, k, blockIdx.x);
    if (__hip_atomic_load(&overflow[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;  // decided already
    for (uint32_t i = threadIdx.x; i < kTileSlots; i += kPalBlock) s_keys[i] = 0;
    if (threadIdx.x == 0) s_count = s_full = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)lo[c] + (uint64_t)(blockIdx.x - tile_start[c]) * kScanTile;
    const uint64_t end = lo[c + 1];
    uint32_t last = 0;
    for (uint32_t q = 0; q < kScanPer; q++) {
        const uint64_t i = base + q * kPalBlock + threadIdx.x;
        if (i >= end) break;
        const uint32_t v = attrs[i];
        if (v == last) continue;
        last = v;
        if (v == 0xFFFFFFFFu) {
            s_full = 1;
            continue;
        }
        uint32_t h = slot_hash(v) & (kTileSlots - 1);
        for (uint32_t probe = 0;; probe++) {
            if (probe == kTileSlots) {
                s_full = 1;
                break;
            }
            uint32_t cur = s_keys[h];
            if (cur == 0) {
                cur = atomicCAS(&s_keys[h], 0u, v);
                if (cur == 0) {  // inserted here
                    if (atomicAdd(&s_count, 1u) >= kPalEntries) s_full = 1;
                    break;
                }
            }
            if (cur == v) break;
            h = (h + 1) & (kTileSlots - 1);
        }
    }
    __syncthreads();
    if (s_full) {
        if (threadIdx.x == 0) atomicOr(&overflow[c], 1u);
        return;
    }
    // merge this tile's values into the chunk's table
    if (threadIdx.x < kTileSlots) {
        const uint32_t v = s_keys[threadIdx.x];
        if (v != 0) {
            uint32_t* t = table + (uint64_t)c * kChunkSlots;
            uint32_t h = slot_hash(v) & (kChunkSlots - 1);
            for (uint32_t probe = 0;; probe++) {
                if (probe == kChunkSlots) {
                    atomicOr(&overflow[c], 1u);
                    break;
                }
                uint32_t cur = __hip_atomic_load(&t[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == 0) {
                    cur = atomicCAS(&t[h], 0u, v);
                    if (cur == 0) {
                        if (atomicAdd(&count[c], 1u) >= kPalEntries) atomicOr(&overflow[c], 1u);
                        break;
                    }
                }
                if (cur == v) break;
                h = (h + 1) & (kChunkSlots - 1);
            }
        }
    }
}

__global__ void __launch_bounds__(kPalBlock) k_palette_final(const uint32_t* table, const uint32_t* count,
                                                              const uint32_t* overflow, uint32_t k,
                                                              uint32_t* palettes, uint32_t* counts_out) {
    const uint32_t c = blockIdx.x * kPalBlock + threadIdx.x;
    if (c >= k) return;
    if (overflow[c] || count[c] > kPalEntries) {
        counts_out[c] = kPaletteOverflow;
        return;
    }
    uint32_t v[kPalEntries];
    uint32_t n = 0;
    for (uint32_t s = 0; s < kChunkSlots; s++) {
        const uint32_t x = table[(uint64_t)c * kChunkSlots + s];
        if (x == 0 || n == kPalEntries) continue;
        // insertion into ascending order
        uint32_t j = n++;
        while (j > 0 && v[j - 1] > x) {
            v[j] = v[j - 1];
            j--;
        }
        v[j] = x;
    }
    for (uint32_t j = 0; j < kPalEntries; j++) palettes[(uint64_t)c * kPalEntries + j] = j < n ? v[j] : (n ? v[0] : 0u);
    counts_out[c] = n;
}

__global__ void __launch_bounds__(kPalBlock) k_palette_pack(const uint32_t* attrs, const uint32_t* lo,
                                                             const uint32_t* tile_start, uint32_t k,
                                                             const uint32_t* counts, const uint32_t* palettes,
                                                             const uint64_t* dst, uint32_t* pool) {
    const uint32_t c = chunk_of(tile_start, k, blockIdx.x);
    const uint32_t tile = blockIdx.x - tile_start[c];
    const uint64_t first = lo[c], n = (uint64_t)lo[c + 1] - first;
    const uint64_t i0 = (uint64_t)tile * kPackTile + (uint64_t)threadIdx.x * 8;
    uint32_t* out = pool + dst[c];
    if (counts[c] == kPaletteOverflow) {
        for (uint32_t j = 0; j < 8; j++) {
            const uint64_t i = (uint64_t)tile * kPackTile + j * kPalBlock + threadIdx.x;  // coalesced
            if (i < n) out[i] = attrs[first + i];
        }
        return;
    }
    uint32_t pal[kPalEntries];
#pragma unroll
    for (uint32_t q = 0; q < kPalEntries; q++) pal[q] = palettes[(uint64_t)c * kPalEntries + q];
    if (tile == 0 && threadIdx.x < kPalEntries) out[threadIdx.x] = pal[threadIdx.x];
    if (i0 >= n) return;
    uint32_t word = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
        if (i0 + j < n) {
            const uint32_t v = attrs[first + i0 + j];
            // the first matching entry (padding repeats entry 0)
            uint32_t idx = 0;
#pragma unroll
            for (int q = kPalEntries - 1; q >= 0; q--) idx = pal[q] == v ? (uint32_t)q : idx;
            word |= idx << (4 * j);
        }
    }
    out[kPalEntries + i0 / 8] = word;
}

template <class T>
PyrResult upload(const std::vector<T>& h, T** d, hipStream_t s) {
    if (pool_alloc((void**)d, std::max<size_t>(h.size() * sizeof(T), 16), s) != hipSuccess) {
        (void)hipGetLastError();
        *d = nullptr;
        return fail(PYR_ERROR_OUT_OF_MEMORY, "palette tables");
    }
    if (hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_CUDA, "palette upload");
    }
    return PYR_OK;
}

}  // namespace

uint64_t palette_words(uint32_t count, uint64_t voxels) {
    return count == kPaletteOverflow ? voxels : kPalEntries + (voxels + 7) / 8;
}

PyrResult palette_scan(const uint32_t* attrs, const std::vector<uint32_t>& lo, hipStream_t s,
                       std::vector<uint32_t>* counts, PaletteScratch* scratch) {
    const uint32_t k = (uint32_t)lo.size() - 1;
    counts->assign(k, kPaletteOverflow);
    *scratch = PaletteScratch();
    if (k == 0) return PYR_OK;
    for (uint32_t j = 0; j < k; j++)  // tiles never run past a chunk's end
        if (lo[j + 1] < lo[j]) return fail(PYR_ERROR_INVALID_ARGUMENT, "chunk %u: attribute ranks decrease", j);
    std::vector<uint32_t> tiles(k + 1, 0);
    for (uint32_t j = 0; j < k; j++) tiles[j + 1] = tiles[j] + (uint32_t)((lo[j + 1] - lo[j] + kScanTile - 1) / kScanTile);
    PyrResult r = upload(lo, &scratch->lo, s);
    uint32_t* d_tiles = nullptr;
    if (r == PYR_OK) r = upload(tiles, &d_tiles, s);
    uint32_t *table = nullptr, *cnt = nullptr, *d_counts = nullptr;
    const uint64_t words = (uint64_t)k * (kChunkSlots + 2) + k;
    if (r == PYR_OK && (pool_alloc((void**)&table, words * 4, s) != hipSuccess ||
                        pool_alloc((void**)&scratch->palettes, (uint64_t)k * kPalEntries * 4, s) != hipSuccess)) {
        (void)hipGetLastError();
        r = fail(PYR_ERROR_OUT_OF_MEMORY, "palette tables");
    }
    if (r == PYR_OK) {
        cnt = table + (uint64_t)k * kChunkSlots;  // [k] counts, [k] overflow flags, [k] results
        d_counts = cnt + 2 * (uint64_t)k;
        if (hipMemsetAsync(table, 0, words * 4, s) != hipSuccess) r = fail(PYR_ERROR_CUDA, "palette clear");
    }
    if (r == PYR_OK && tiles[k]) {
        hipLaunchKernelGGL(k_palette_scan, dim3(tiles[k]), dim3(kPalBlock), 0, s, attrs, scratch->lo, d_tiles, k,
                           table, cnt, cnt + k);
        hipLaunchKernelGGL(k_palette_final, dim3((k + kPalBlock - 1) / kPalBlock), dim3(kPalBlock), 0, s, table, cnt,
                           cnt + k, k, scratch->palettes, d_counts);
        if (hipMemcpyAsync(counts->data(), d_counts, (uint64_t)k * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            (void)hipGetLastError();
            r = fail(PYR_ERROR_CUDA, "palette scan");
        }
    }
    pool_free(d_tiles, s);
    if (r == PYR_OK) {
        // keep the device counts for the pack step (they sit in the table block)
        scratch->table = table;
        scratch->counts = d_counts;
    } else {
        pool_free(table, s);
    }
    return r;
}

PyrResult palette_pack(const uint32_t* attrs, const std::vector<uint32_t>& lo, const std::vector<uint64_t>& dst,
                       uint32_t* pool, hipStream_t s, PaletteScratch* scratch) {
    const uint32_t k = (uint32_t)lo.size() - 1;
    if (k == 0) return PYR_OK;
    std::vector<uint32_t> tiles(k + 1, 0);
    for (uint32_t j = 0; j < k; j++) tiles[j + 1] = tiles[j] + (uint32_t)((lo[j + 1] - lo[j] + kPackTile - 1) / kPackTile);
    uint32_t* d_tiles = nullptr;
    uint64_t* d_dst = nullptr;
    PyrResult r = upload(tiles, &d_tiles, s);
    if (r == PYR_OK) r = upload(dst, &d_dst, s);
    if (r == PYR_OK && tiles[k])
        hipLaunchKernelGGL(k_palette_pack, dim3(tiles[k]), dim3(kPalBlock), 0, s, attrs, scratch->lo, d_tiles, k,
                           scratch->counts, scratch->palettes, d_dst, pool);
    // host vectors are read by the copies above: finish before returning
    if (r == PYR_OK && hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        r = fail(PYR_ERROR_CUDA, "palette pack");
    }
    pool_free(d_tiles, s);
    pool_free(d_dst, s);
    return r;
}

void palette_release(PaletteScratch* scratch, hipStream_t s) {
    pool_free(scratch->lo, s);
    pool_free(scratch->palettes, s);
    pool_free(scratch->table, s);
    *scratch = PaletteScratch();
}

}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_palette_pack(uint64_t attrs, const uint32_t* first, uint32_t count, uint64_t out,
                                      uint64_t out_words, uint64_t* word_offset, uint32_t* palette_count,
                                      void* stream) {
    diag_clear();
    if (!first || !word_offset || !palette_count || (!attrs && count && first[count]))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    std::vector<uint32_t> lo(first, first + count + 1);
    for (uint32_t j = 0; j < count; j++)
        if (lo[j + 1] < lo[j]) return fail(PYR_ERROR_INVALID_ARGUMENT, "first[] must not decrease (chunk %u)", j);
    hipStream_t s = (hipStream_t)stream;
    std::vector<uint32_t> counts;
    PaletteScratch ps;
    PyrResult r = palette_scan((const uint32_t*)(uintptr_t)attrs, lo, s, &counts, &ps);
    if (r == PYR_OK) {
        uint64_t words = 0;
        for (uint32_t j = 0; j < count; j++) {
            word_offset[j] = words;
            palette_count[j] = counts[j];
            words += palette_words(counts[j], lo[j + 1] - lo[j]);
        }
        word_offset[count] = words;
        if (out && out_words < words)
            r = fail(PYR_ERROR_INVALID_ARGUMENT, "output of %llu words, %llu needed", (unsigned long long)out_words,
                     (unsigned long long)words);
        else if (out) {
            std::vector<uint64_t> dst(word_offset, word_offset + count);
            r = palette_pack((const uint32_t*)(uintptr_t)attrs, lo, dst, (uint32_t*)(uintptr_t)out, s, &ps);
        }
    }
    palette_release(&ps, s);
    (void)hipStreamSynchronize(s);
    return r;
}

}  // extern "C"
