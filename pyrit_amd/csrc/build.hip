// build.hip - GPU sparse-voxel-DAG build for gfx950.
//
// Produces exactly the arrays the reference GPU builder produces
// (src/device/gbuild.zig, op order of src/gpu_build.zig:135-360): nodes level by
// level bottom-up with absolute word refs, each level's nodes in the order of
// their first appearance in Morton order, leaves in first-appearance order,
// attributes in DFS rank order. The pipeline is this file's own:
//
//   encode      PyrVoxel -> (key, attr) + LDS histograms of every radix digit
//   sort        P onesweep LSD passes: in-wave ranking by LDS lane masks (and
//               ballots for the first digits of a row), per-wave digit
//               counters in LDS, decoupled look-back per digit, LDS-staged
//               coalesced scatter; with the bucket finish only the key bits
//               from 14 up are sorted
//   unique +    either one fused tile scan over the sorted records (last
//   bricks      record of a key wins, attribute 0 deletes, brick heads, masks
//               by wave-segmented OR) or the bucket finish: one LDS key table
//               per bucket of 2^14 keys (atomic max of the record index)
//   leaves      open-addressing HBM hash, 64-bit CAS, owner = atomicMin index,
//               compaction by tile scan
//   levels      per level: parent tuples (tile scan), tuple hash (CAS of
//               index + 1, word compare), owner = atomicMin, compaction of the
//               owners' words by tile scan, pointer remap via the owners
//   finalize    roots, empty geometry
//
// Every count lives on the device; kernels are launched with host upper bounds
// and read the true counts, so one build is a fixed launch sequence with a
// single read-back at the end (graph-capturable: pyrhip_graph_*).
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "common.h"
#include "internal.h"

namespace pyr {

// ---------------------------------------------------------------------------
// counters
// ---------------------------------------------------------------------------
enum : uint32_t {
    C_ERR = 0,     // non-zero: a bounded wait gave up or a table filled up
    C_U = 1,       // unique voxels
    C_B = 2,       // bricks
    C_LEAVES = 3,  // distinct leaves
    C_ROOT = 4,
    C_WORDS = 5,
    C_TILES = 6,   // record tiles of a batch (segment-aligned)
    C_NCH = 8,     // + lv: elements of level lv (children of the level-lv parents)
    C_CUR = 40,    // + lv: node-word cursor when level lv's parents are written
    C_PASS = 64,   // bit p: radix pass p moves records (some segment has two digits)
    C_PSRC = 65,   // bit p: record buffer pass p reads (it writes the other)
    C_FINAL = 66,  // record buffer that holds the sorted records
    C_BKT = 67,    // buckets of a batch's bucket finish
    C_TILE = 72,   // tile counters, one per launch
    C_COUNT = 256,
};
constexpr uint32_t kMaxLaunches = C_COUNT - C_TILE;

constexpr int kScanItems = 8;                     // elements per thread in a scan tile
constexpr int kScanTile = kBlock * kScanItems;    // 2048
constexpr int kNodeStride = 10;                   // words per parent in node_tmp (gbuild.zig:66)

// table capacity for n keys at a maximum load of load_pct percent: the next
// power of two >= n * 100 / load_pct, at least 1024 (80 %: gpu_build.zig:275-277)
__host__ __device__ inline uint64_t hash_cap_for(uint64_t n, uint32_t load_pct) {
    uint64_t want = (n * 100 + load_pct - 1) / load_pct, c = 1024;
    while (c < want) c <<= 1;
    return c;
}

// ---------------------------------------------------------------------------
// record segments and tiles
//
// The record array is cut into segments: chunk c of a batch owns positions
// [offsets[c], offsets[c+1]), a single build is one segment [0, n). Records
// are sorted by their chunk-local key inside their segment only, so the chunk
// index never enters the sort key (fewer digits, and local keys of up to 32
// bits travel packed with their attribute). Tiles of `tile` records never
// cross a segment boundary; a batch describes its tiles in device memory
// ({segment, first position, length, first tile of the segment}), a single
// build derives them from the tile index.
// ---------------------------------------------------------------------------
struct Segs {
    uint32_t single;          // 1: one segment [0, n), no descriptors
    uint32_t count;           // segments
    uint32_t n;               // records
    uint32_t tile;            // records per tile
    const uint32_t* off;      // [count + 1] segment starts (batches)
    const uint4* desc;        // [tiles] (batches)
    const uint32_t* tiles;    // device tile count (batches)
};

__device__ __forceinline__ uint32_t tile_count(const Segs& g) {
    return g.single ? (g.n + g.tile - 1) / g.tile : *g.tiles;
}

__device__ __forceinline__ uint4 tile_desc(const Segs& g, uint32_t t) {
    if (g.single) {
        const uint32_t start = t * g.tile;
        return make_uint4(0u, start, min(g.tile, g.n - start), 0u);
    }
    return g.desc[t];
}

__device__ __forceinline__ uint32_t seg_end(const Segs& g, uint32_t seg) { return g.single ? g.n : g.off[seg + 1]; }

// tile_first[s] = tiles of segments before s (one block; segments are
// scanned 256 at a time), tile_first[count] = *total = the tile count.
// Offsets that are not a partition of [0, n) (of [0, m) with m <= n when
// `at_most`: a graph sized for n records replays smaller batches) raise error
// bit 32 and leave no tiles, so no later kernel touches a record.
__global__ void __launch_bounds__(kBlock) k_seg_tiles(const uint32_t* off, uint32_t count, uint32_t n,
                                                       uint32_t at_most, uint32_t tile, uint32_t* tile_first,
                                                       uint32_t* total, uint32_t* err) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    __shared__ uint32_t s_bad;
    if (threadIdx.x == 0) s_bad = off[0] != 0 || (at_most ? off[count] > n : off[count] != n);
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t b = 0; b < count; b += kBlock) {
        const uint32_t s = b + threadIdx.x;
        uint32_t len = 0;
        if (s < count) {
            const uint32_t lo = off[s], hi = off[s + 1];
            if (hi < lo) s_bad = 1;
            len = hi >= lo ? hi - lo : 0;
        }
        uint32_t v[1] = {(uint32_t)(((uint64_t)len + tile - 1) / tile)}, e[1];
        const uint32_t sum = tile_scan<1>(v, e, s_wave);
        if (s < count) tile_first[s] = carry + e[0];
        carry += sum;
        __syncthreads();  // s_wave is reused by the next round
    }
    if (threadIdx.x == 0) {
        tile_first[count] = carry;
        *total = s_bad ? 0u : carry;
        if (s_bad) atomicOr(err, 32u);
    }
}

// tile descriptors, one wave per segment (nothing is written for offsets
// that k_seg_tiles rejected)
__global__ void __launch_bounds__(kBlock) k_seg_desc(const uint32_t* off, const uint32_t* tile_first, uint32_t count,
                                                      uint32_t tile, const uint32_t* total, uint4* desc) {
    if (*total == 0) return;
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t s = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; s < count; s += waves) {
        const uint32_t first = tile_first[s], nt = tile_first[s + 1] - first;
        const uint32_t start = off[s], len = off[s + 1] - start;
        for (uint32_t t = lane_id(); t < nt; t += kWave)
            desc[first + t] = make_uint4(s, start + t * tile, min(tile, len - t * tile), first);
    }
}

// ---------------------------------------------------------------------------
// input description shared by the encode kernel
// ---------------------------------------------------------------------------
struct EncodeArgs {
    const uint64_t* old_keys;
    const uint32_t* old_vals;
    uint32_t n_old;
    const int4* voxels;       // PyrVoxel records
    uint32_t n_new;
    uint32_t log2_size;
    // chunk batches
    uint32_t chunk_count;     // 0 = single build
    uint32_t capacity;        // 0 = packed segments
    const uint32_t* offsets;  // [chunk_count + 1], offsets[0] = 0
    const uint8_t* chunk_log2;
    uint32_t chunk_bits;
    // outputs
    uint64_t* keys;
    uint32_t* vals;
    uint32_t* hist;           // [passes][segments][radix]
    uint32_t passes, radix_bits;
    uint32_t shift0;          // first key bit of the lowest digit (bucket finish: bits below are not sorted)
    uint32_t packed;          // keys[i] = key << 32 | attribute, vals unused
    // pass-0 digit counts per tile (8-bit digits; null: not counted): for
    // tile t of block b's run, tile_run[t][d] = records of digit d in the
    // run's tiles before t; run_sum[b][d] = the whole run's
    uint32_t* tile_run;
    uint32_t* run_sum;
};

// Input record i of the record array (the `local`-th record of segment
// `seg`), or x = -1 (out of range) when it has none: a batch record past its
// chunk's capacity. Old keys of an edit build are not read here.
__device__ __forceinline__ int4 fetch_voxel(const EncodeArgs& a, uint32_t seg, uint32_t local, uint32_t i) {
    if (i < a.n_old || (a.capacity && local >= a.capacity)) return make_int4(-1, -1, -1, 0);
    const uint64_t src = a.chunk_count == 0 ? (uint64_t)(i - a.n_old)
                                            : (a.capacity ? (uint64_t)seg * a.capacity + local : (uint64_t)i);
    // (every input record is read once: no reason to keep it in the caches)
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i r = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(a.voxels) + src);
    return make_int4(r.x, r.y, r.z, r.w);
}

// Chunk-local sort key and attribute of record i from its fetched voxel;
// out-of-range records get the invalid bit just above the local key bits, so
// they sort to the end of their segment. encode (gbuild.zig:184-199) and
// pack_chunks (gbuild.zig:203-230) without the chunk index in the key.
// Edge of segment seg's chunk (a chunk size above the batch's counts as the
// batch's: keys stay below the invalid bit).
__device__ __forceinline__ int32_t chunk_edge(const EncodeArgs& a, uint32_t seg) {
    return 1 << (a.chunk_count && a.chunk_log2 ? min((uint32_t)a.chunk_log2[seg], a.log2_size) : a.log2_size);
}

// (n: the edge of the record's chunk, chunk_edge)
__device__ __forceinline__ void encode_one(const EncodeArgs& a, int32_t n, uint32_t i, int4 v, uint64_t& key,
                                           uint32_t& val) {
    const uint32_t lbits = local_bits(a.log2_size);
    if (i < a.n_old) {  // earlier build's sorted unique voxels (single builds)
        key = a.old_keys[i];
        val = a.old_vals[i];
        return;
    }
    key = 1ull << lbits;
    val = 0;
    if (v.x >= 0 && v.y >= 0 && v.z >= 0 && v.x < n && v.y < n && v.z < n) {
        key = local_key((uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z);
        val = (uint32_t)v.w;
    }
}

constexpr int kEncodeUnroll = 4;  // records per thread whose loads are in flight together

// Records of a contiguous run of tiles per block, plus the digit histograms
// of every pass per segment: counted in LDS and flushed to the global
// histogram (one atomic per non-zero bin) whenever the block's run moves to
// the next segment.
// Writes record i and counts its digits of every pass in the LDS histogram.
// Neighbouring input records often share their high digits: a digit common
// to every active lane of the wave costs one LDS atomic instead of up to 64
// serialised ones on the same bin. (Matching up to four digits of a wave
// this way made the 1024^3 encode 19 % slower, and a second digit alone 3 %:
// the matching costs more VALU work per record than the serialised atomics.)
__device__ __forceinline__ void emit_record(const EncodeArgs& a, uint32_t* s_hist, uint32_t radix, uint32_t i,
                                            uint64_t key, uint32_t val) {
    if (a.packed) {
        __builtin_nontemporal_store((key << 32) | val, &a.keys[i]);
    } else {
        a.keys[i] = key;
        a.vals[i] = val;
    }
    const uint64_t active = __ballot(1);
    for (uint32_t p = 0; p < a.passes; p++) {
        const uint32_t dg = (uint32_t)((key >> (a.shift0 + p * a.radix_bits)) & (radix - 1));
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(dg);
        const uint64_t same = __ballot(dg == d0);
        if (same == active) {
            if (lane_id() == (uint32_t)__ffsll((unsigned long long)active) - 1)
                atomicAdd(&s_hist[p * radix + d0], (uint32_t)__popcll(same));
        } else {
            atomicAdd(&s_hist[p * radix + dg], 1u);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_encode(EncodeArgs a, Segs g, uint32_t tiles_per_block) {
    extern __shared__ uint32_t s_hist[];
    const uint32_t radix = 1u << a.radix_bits, bins = a.passes * radix;
    for (uint32_t i = threadIdx.x; i < bins; i += kBlock) s_hist[i] = 0;
    const uint32_t tiles = tile_count(g);
    const uint32_t t0 = blockIdx.x * tiles_per_block;
    const uint32_t t1 = min(tiles, t0 + tiles_per_block);
    uint32_t cur = 0xFFFFFFFFu;
    // counted pass 0 (radix 256: thread = digit): the run's count of the
    // thread's digit so far, and the LDS bin as last seen (bins are cleared
    // when the run enters the next segment)
    const bool counted = a.tile_run != nullptr;
    uint32_t run_acc = 0, seen = 0;
    auto flush = [&]() {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < bins; b += kBlock) {
            const uint32_t c = s_hist[b];
            if (c) {
                const uint32_t p = b / radix, d = b % radix;
                atomicAdd(&a.hist[((uint64_t)p * g.count + cur) * radix + d], c);
                s_hist[b] = 0;
            }
        }
        seen = 0;
        __syncthreads();
    };
    __syncthreads();
    for (uint32_t t = t0; t < t1; t++) {
        const uint4 d = tile_desc(g, t);
        if (d.x != cur) {
            if (cur != 0xFFFFFFFFu) flush();
            cur = d.x;
        }
        if (counted) a.tile_run[(uint64_t)t * kBlock + threadIdx.x] = run_acc;
        const uint32_t local0 = (t - d.w) * g.tile;
        const int32_t edge = chunk_edge(a, d.x);
        for (uint32_t j0 = 0; j0 < d.z; j0 += kBlock * kEncodeUnroll) {
            // the input loads of kEncodeUnroll records are issued before any is used
            int4 raw[kEncodeUnroll];
#pragma unroll
            for (int u = 0; u < kEncodeUnroll; u++) {
                const uint32_t j = j0 + u * kBlock + threadIdx.x;
                raw[u] = j < d.z ? fetch_voxel(a, d.x, local0 + j, d.y + j) : make_int4(-1, -1, -1, 0);
            }
#pragma unroll
            for (int u = 0; u < kEncodeUnroll; u++) {
                const uint32_t j = j0 + u * kBlock + threadIdx.x;
                if (j < d.z) {
                    uint64_t key;
                    uint32_t val;
                    encode_one(a, edge, d.y + j, raw[u], key, val);
                    emit_record(a, s_hist, radix, d.y + j, key, val);
                }
            }
        }
        if (counted) {
            __syncthreads();
            const uint32_t v = s_hist[threadIdx.x];
            run_acc += v - seen;
            seen = v;
            __syncthreads();  // every bin read before the next tile counts into it
        }
    }
    if (cur != 0xFFFFFFFFu) flush();
    if (counted) a.run_sum[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = run_acc;
}

// Exclusive prefix of the runs' digit counts over the runs, per digit:
// run_off[b][d] = records of digit d in runs before b. One block per digit;
// thread t sums a contiguous group of runs, a block scan of the groups, then
// every thread writes its group's prefixes.
__global__ void __launch_bounds__(kBlock) k_run_scan(const uint32_t* run_sum, uint32_t* run_off, uint32_t runs) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    const uint32_t d = blockIdx.x, per = (runs + kBlock - 1) / kBlock;
    const uint32_t lo = min(runs, threadIdx.x * per), hi = min(runs, lo + per);
    uint32_t v[1] = {0}, e[1];
    for (uint32_t b = lo; b < hi; b++) v[0] += run_sum[(uint64_t)b * kBlock + d];
    tile_scan<1>(v, e, s_wave);
    uint32_t acc = e[0];
    for (uint32_t b = lo; b < hi; b++) {
        const uint32_t c = run_sum[(uint64_t)b * kBlock + d];
        run_off[(uint64_t)b * kBlock + d] = acc;
        acc += c;
    }
}

// Per (pass, segment): exclusive scan of the digit counts plus the segment's
// start, the first output position of each digit in that pass.
// A pass in which every record of a segment has the same digit leaves that
// segment as it is; passes that move no record in any segment are marked
// skippable in counters[C_PASS] (one bit per pass that moves records).
__global__ void __launch_bounds__(kBlock) k_digit_scan(const uint32_t* hist, uint32_t* base, Segs g,
                                                       uint32_t passes, uint32_t radix, uint32_t* counters) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    const uint32_t seg = blockIdx.x % g.count, pass = blockIdx.x / g.count;
    const uint64_t row = (uint64_t)blockIdx.x * radix;  // blockIdx.x = pass * count + seg
    const uint32_t per = radix / kBlock;
    uint32_t v[1] = {0}, e[1];
    for (uint32_t k = 0; k < per; k++) v[0] += hist[row + threadIdx.x * per + k];
    const uint32_t len = tile_scan<1>(v, e, s_wave);
    uint32_t acc = (g.single ? 0u : g.off[seg]) + e[0];
    bool whole = false;  // one digit holds the whole segment
    for (uint32_t k = 0; k < per; k++) {
        const uint64_t d = row + threadIdx.x * per + k;
        base[d] = acc;
        acc += hist[d];
        whole |= hist[d] == len;
    }
    const int stays = __syncthreads_or(whole ? 1 : 0);
    if (threadIdx.x == 0 && len > 0 && !stays) atomicOr(&counters[C_PASS], 1u << pass);
}

// Source buffer of every radix pass and the buffer that ends up sorted:
// skipped passes do not swap the buffers.
__global__ void k_pass_plan(uint32_t passes, uint32_t* counters) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t moves = counters[C_PASS];
    uint32_t cur = 0, src = 0;
    for (uint32_t p = 0; p < passes; p++)
        if ((moves >> p) & 1) {
            src |= cur << p;
            cur ^= 1;
        }
    counters[C_PSRC] = src;
    counters[C_FINAL] = cur;
}

// ---------------------------------------------------------------------------
// onesweep LSD pass
// ---------------------------------------------------------------------------
// the two record buffers of the ping-pong (keys, and attributes when not packed)
struct RecordBufs {
    uint64_t* keys[2];
    uint32_t* vals[2];
};

struct SortArgs {
    RecordBufs b;
    uint32_t pass;
    const uint32_t* plan;  // counters: C_PASS, C_PSRC
    Segs g;
    uint32_t shift;
    const uint32_t* digit_base;  // [segments][radix] first position of each digit in each segment, this pass
    uint64_t* status;            // [tiles][radix], wide words (epoch-tagged)
    uint32_t* status32;          // [tiles][radix], narrow words (cleared before the pass)
    uint32_t epoch;
    uint32_t* tile_ctr;
    uint32_t* err;
    uint32_t peel;  // digits matched by ballots before the LDS match, per row
    // counted pass (pass 0 over records in input order, 8-bit digits): the
    // tile's digit offsets come from the encode counts, no look-back
    const uint32_t* tile_run;
    const uint32_t* run_off;
    uint32_t run_tiles;  // tiles per encode run
};

// The passes of the bucket finish sort the high key bits only, which
// neighbouring input records mostly share: a row of 64 records then holds a
// few digits, whose lanes would all OR into the same LDS word (serialised).
// Up to this many digits of a row are matched with one ballot each first.
// (On the 1024^3 heightfield: 3.47 -> 3.05 ms per pass; the uniform low
// digits of batch passes gain nothing from it, so those keep 0.)
constexpr uint32_t kSortPeel = 4;

// predecessor status words a sort tile's digit thread loads in the first
// look-back round and in every later one (1024^3 heightfield, both passes:
// 1/16 4.99 ms, 2/8 4.64, 4/16 4.94, 8/8 4.58, 16/16 4.70)
constexpr int kPollFirst = 8;
constexpr int kPoll = 8;

// PACKED: one u64 word per record, key << 32 | attribute (keys of at most 32
// bits); the digit is taken at a.shift >= 32 and there is no value stream.
// NARROW: look-back status words of 32 bits, [31:30] flag | [29:0] value,
// for segments of fewer than 2^30 records; the host clears them before the
// pass (no epoch tag). The look-back reads of a tile are most of what a
// pass reads beyond its records, so halving the word halves that traffic.
template <int RB, int ITEMS, bool PACKED, bool NARROW>
__global__ void __launch_bounds__(kBlock) k_sort_pass(SortArgs a) {
    // tile of kBlock * ITEMS keys (the names shadow the defaults above)
    constexpr int kSortItems = ITEMS;
    constexpr int kSortTile = kBlock * ITEMS;
    constexpr uint32_t R = 1u << RB;
    constexpr uint32_t DPT = R / kBlock;  // digits per thread
    constexpr int W = kBlock / kWave;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_cnt[W][R];  // per-wave digit counters, then wave-exclusive prefixes
    __shared__ uint32_t s_local[R];        // tile-local exclusive digit offsets
    __shared__ uint32_t s_part[kBlock];
    __shared__ uint64_t s_dst[R];          // global position of local slot 0 per digit
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[PACKED ? 1 : kSortTile];

    if (!((a.plan[C_PASS] >> a.pass) & 1)) return;  // no segment has two digits: nothing moves
    const uint32_t src = (a.plan[C_PSRC] >> a.pass) & 1;
    const uint64_t* keys_in = a.b.keys[src];
    const uint32_t* vals_in = a.b.vals[src];
    uint64_t* keys_out = a.b.keys[src ^ 1];
    uint32_t* vals_out = a.b.vals[src ^ 1];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
    // per-wave lane masks of each digit while ranking one row (zero between
    // rows); the staging area is free until the ranking is done
    // (wider digits whose masks do not fit match the digit with RB ballots)
    constexpr bool kLdsMatch = W * R <= kSortTile;
    uint64_t* s_match = s_keys;
    if (tid == 0) s_tile = atomicAdd(a.tile_ctr, 1u);
    for (uint32_t i = tid; i < W * R; i += kBlock) {
        (&s_wave_cnt[0][0])[i] = 0;
        if constexpr (kLdsMatch) s_match[i] = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= tile_count(a.g)) return;  // grids are sized by an upper bound of the tile count
    // tile of segment td.x: records [td.y, td.y + td.z); td.w = the segment's first tile
    const uint4 td = tile_desc(a.g, tile);
    const uint32_t count = td.z;
    const uint32_t* seg_base = a.digit_base + (uint64_t)td.x * R;

    uint64_t k[kSortItems];
    uint32_t v[kSortItems], rank[kSortItems];
    const uint32_t wbase = wave * kWave * kSortItems;
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        const uint32_t li = wbase + i * kWave + lane;
        bool ok = li < count;
        k[i] = ok ? __builtin_nontemporal_load(&keys_in[td.y + li]) : 0;
        if constexpr (!PACKED) v[i] = ok ? vals_in[td.y + li] : 0;
    }
    // stable ranking inside the wave: up to a.peel digits of the row are
    // matched by ballots (the first remaining lane's digit each time); the
    // other lanes OR their lane bit into the wave's mask of their digit (one
    // LDS atomic; lanes of one digit conflict, others do not), read it back
    // and clear it; the lowest peer bumps the wave's counter for that digit. (Matching the
    // digit with RB ballots costs about nine VALU operations per bit.)
    // LDS accesses are relaxed workgroup-scope atomics: the wave's lanes
    // execute each LDS instruction together and in program order, and
    // same-address atomics keep that order through the compiler (volatile
    // would turn them into flat accesses that wait on the vector memory
    // counter)
    const uint64_t lt = (1ull << lane) - 1;
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        bool ok = wbase + i * kWave + lane < count;
        uint32_t d = (uint32_t)(k[i] >> a.shift) & (R - 1);
        uint64_t peers = 0;
        if constexpr (kLdsMatch) {
            uint64_t rest = __ballot(ok);
            for (uint32_t t = 0; t < a.peel && rest; t++) {
                const uint32_t d0 = __builtin_amdgcn_readlane(d, (int)__ffsll((unsigned long long)rest) - 1);
                const uint64_t m = __ballot(d == d0) & rest;
                if (d == d0) peers = m;
                rest &= ~m;
            }
            const bool lds = ok && ((rest >> lane) & 1);
            uint64_t* mp = &s_match[wave * R + d];
            if (rest) {
                if (lds) __hip_atomic_fetch_or(mp, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (lds) peers = __hip_atomic_load(mp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (lds) __hip_atomic_store(mp, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        } else {
            peers = __ballot(ok);
#pragma unroll
            for (int b = 0; b < RB; b++) {
                uint64_t m = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? m : ~m;
            }
        }
        uint32_t* cp = &s_wave_cnt[wave][d];
        const uint32_t c = __hip_atomic_load(cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        uint32_t below = (uint32_t)__popcll(peers & lt);
        rank[i] = c + below;
        if (ok && below == 0)
            __hip_atomic_store(cp, c + (uint32_t)__popcll(peers), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    // per digit: wave-exclusive prefixes and the tile total
    uint32_t tot[DPT];
    uint32_t tsum = 0;
#pragma unroll
    for (uint32_t j = 0; j < DPT; j++) {
        const uint32_t d = tid * DPT + j;
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            uint32_t c = s_wave_cnt[w][d];
            s_wave_cnt[w][d] = acc;
            acc += c;
        }
        tot[j] = acc;
        tsum += acc;
    }
    // tile-local exclusive offsets of the digits (thread t owns digits t*DPT..)
    s_part[tid] = tsum;
    __syncthreads();
    if (wave == 0) {
        uint32_t x0 = s_part[lane * 4 + 0], x1 = s_part[lane * 4 + 1], x2 = s_part[lane * 4 + 2],
                 x3 = s_part[lane * 4 + 3];
        uint32_t s = x0 + x1 + x2 + x3;
        uint32_t inc = wave_inclusive_scan(s);
        uint32_t e = inc - s;
        s_part[lane * 4 + 0] = e;
        s_part[lane * 4 + 1] = e + x0;
        s_part[lane * 4 + 2] = e + x0 + x1;
        s_part[lane * 4 + 3] = e + x0 + x1 + x2;
    }
    __syncthreads();
    {
        uint32_t acc = s_part[tid];
#pragma unroll
        for (uint32_t j = 0; j < DPT; j++) {
            s_local[tid * DPT + j] = acc;
            acc += tot[j];
        }
    }
    // decoupled look-back per digit over the tiles of the segment: publish
    // the tile count, then the prefix
    const int64_t first = td.w;
    using SW = typename std::conditional<NARROW, uint32_t, uint64_t>::type;
    // status word of flag 1 (tile count) or 2 (inclusive prefix), its flag
    // (0 = not published in this pass) and its value
    auto make = [&](uint32_t f, uint64_t v) -> SW {
        if constexpr (NARROW) return (SW)((f << 30) | (uint32_t)v);
        else return status_word((uint64_t)f << 62, a.epoch, v);
    };
    auto flag_of = [&](SW w) -> uint32_t {
        if constexpr (NARROW) return (uint32_t)(w >> 30);
        else return status_flag(w, a.epoch);
    };
    auto value_of = [&](SW w) -> uint64_t {
        if constexpr (NARROW) return w & 0x3FFFFFFFu;
        else return w & kValueMask;
    };
    // publish the tile's counts first, then stage the tile in LDS in digit
    // order (local offsets only) while the predecessors finish, then look back
    // (a counted pass knows its offsets: nothing to publish or poll)
    const bool counted = a.tile_run != nullptr;
#pragma unroll
    for (uint32_t j = 0; j < DPT; j++) {
        const uint32_t d = tid * DPT + j;
        SW* st;
        if constexpr (NARROW) st = a.status32 + d;
        else st = a.status + d;
        if (!counted)
            __hip_atomic_store(&st[(uint64_t)tile * R], make(tile == td.w ? 2 : 1, tot[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();  // s_local is complete
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        if (wbase + i * kWave + lane < count) {
            uint32_t d = (uint32_t)(k[i] >> a.shift) & (R - 1);
            uint32_t pos = s_local[d] + s_wave_cnt[wave][d] + rank[i];
            s_keys[pos] = k[i];
            if constexpr (!PACKED) s_vals[pos] = v[i];
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < DPT; j++) {
        const uint32_t d = tid * DPT + j;
        SW* st;
        if constexpr (NARROW) st = a.status32 + d;
        else st = a.status + d;
        uint64_t excl = 0;
        if (counted) {
            // records of digit d in the segment's tiles before this one: the
            // difference of two prefixes over all tiles (run prefix + the
            // run-local prefix), this tile's and the segment's first tile's
            const uint32_t f = td.w;
            const uint32_t pt = a.run_off[(uint64_t)(tile / a.run_tiles) * R + d] + a.tile_run[(uint64_t)tile * R + d];
            const uint32_t pf = a.run_off[(uint64_t)(f / a.run_tiles) * R + d] + a.tile_run[(uint64_t)f * R + d];
            excl = pt - pf;
        } else if (tile != td.w) {  // not the first tile of its segment
            // predecessors are polled newest to oldest and summed until an
            // inclusive prefix: kPollFirst in the first round (the previous
            // tile's prefix is often there already), kPoll (independent loads
            // in flight) in every later one
            int64_t t = (int64_t)tile - 1;
            Wait wt;
            int width = kPollFirst;
            while (t >= first) {
                SW w[kPoll];
#pragma unroll
                for (int q = 0; q < kPoll; q++)
                    w[q] = q < width && t - q >= first
                               ? __hip_atomic_load(&st[(uint64_t)(t - q) * R], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                               : make(2, 0);  // before the segment: prefix 0
                int used = 0;
                bool done = false, wait = false;
#pragma unroll
                for (int q = 0; q < kPoll; q++) {
                    if (done || wait || q >= width) continue;
                    const uint32_t f = flag_of(w[q]);
                    if (f == 0) {
                        wait = true;
                        continue;
                    }
                    excl += value_of(w[q]);
                    used++;
                    done = f == 2;
                }
                if (done) break;
                t -= used;
                width = kPoll;
                if (wait && !wait_more(wt, a.err, 2u)) break;
            }
            __hip_atomic_store(&st[(uint64_t)tile * R], make(2, excl + tot[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_dst[d] = (uint64_t)seg_base[d] + excl;
    }
    __syncthreads();
    // coalesced write-out: consecutive slots of one digit are consecutive in HBM
    for (uint32_t j = tid; j < count; j += kBlock) {
        uint64_t key = s_keys[j];
        uint32_t d = (uint32_t)(key >> a.shift) & (R - 1);
        uint64_t dst = s_dst[d] + (j - s_local[d]);
        if (dst >= a.g.n) {  // only after a failed look-back
            atomicOr(a.err, 16u);
            continue;
        }
        __builtin_nontemporal_store(key, &keys_out[dst]);
        if constexpr (!PACKED) vals_out[dst] = s_vals[j];
    }
}

// ---------------------------------------------------------------------------
// tile-scan kernels (persistent blocks, tiles handed out in order)
// ---------------------------------------------------------------------------
struct ScanCtx {
    uint64_t* status;
    uint32_t epoch;
    uint32_t* tile_ctr;
    uint32_t* counters;
};

// Returns the tile's global exclusive prefix; every thread gets it. `total`
// is the tile aggregate from tile_scan. Uses s_bcast (LDS, 1 x u64).
__device__ __forceinline__ uint64_t tile_prefix(const ScanCtx& sc, uint32_t tile, uint32_t total,
                                                uint64_t* s_bcast) {
    if (threadIdx.x < kWave) {
        uint64_t p = lookback_wave(sc.status, tile, total, sc.epoch, &sc.counters[C_ERR]);
        if (threadIdx.x == 0) *s_bcast = p;
    }
    __syncthreads();
    return *s_bcast;
}

__device__ __forceinline__ uint32_t next_tile(const ScanCtx& sc, uint32_t* s_tile) {
    __syncthreads();  // previous tile's LDS use is over
    if (threadIdx.x == 0) *s_tile = atomicAdd(sc.tile_ctr, 1u);
    __syncthreads();
    return *s_tile;
}

template <int ITEMS = kScanItems>
__device__ __forceinline__ uint64_t scan_index(uint32_t tile, int i) {
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    return (uint64_t)tile * (kBlock * ITEMS) + (uint64_t)wave * kWave * ITEMS + (uint64_t)i * kWave + lane;
}

// Kept records and brick heads of one segment-aligned scan piece in a single
// pass (gbuild.zig:366-428: uniqueFlags/uniqueScatter, brickHeads/brickBuild).
//
// A record is kept when it is the last of its key in its segment, valid and
// not a delete. A kept record heads its brick when no earlier kept record of
// the segment has the same brick (key >> 6). Records of one brick are
// adjacent in the sorted array, so "earlier kept record of the brick" means
// "kept record earlier in the same brick run": wave ballots give it inside a
// 64-record row, a carry passes it from row to row and from wave to wave, and
// for the run that reaches into the piece from before, wave 0 looks back
// through the preceding records of that run in memory. Two exclusive scans
// (kept, heads) and two decoupled look-backs, run by waves 0 and 1 at the
// same time, give every kept record its unique-voxel position and its brick
// index; bricks' masks are OR-reduced per row and merged with one atomicOr
// per (brick, row).
template <int ITEMS, bool PACKED>
__global__ void __launch_bounds__(kBlock, 4) k_unique_bricks(ScanCtx su, uint64_t* status_heads, Segs g, uint32_t subs,
                                                           RecordBufs b, uint32_t lbits,
                                                           uint64_t* brick_key, uint64_t* masks,
                                                           uint32_t* first_voxel, uint64_t brick_cap) {
    constexpr int W = kBlock / kWave;
    // packed records carry keys of at most 32 bits: compare and move them as such
    using K = typename std::conditional<PACKED, uint32_t, uint64_t>::type;
    constexpr K kNone = (K)~0ull;
    __shared__ uint32_t s_tile, s_wave[W], s_wave2[W];
    __shared__ uint64_t s_pre[2];
    __shared__ uint32_t s_any_start[W], s_any_kept[W], s_out[W], s_carry0;
    const K invalid = (K)(1ull << lbits);
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    const uint64_t lt = (1ull << lane) - 1;
    const uint32_t fin = su.counters[C_FINAL];
    const uint64_t* keys = b.keys[fin];
    const uint32_t* vals = b.vals[fin];
    uint64_t* ukey = b.keys[fin ^ 1];
    uint32_t* uval = PACKED ? b.vals[0] : b.vals[fin ^ 1];
    const uint32_t ntiles = tile_count(g) * subs;
    auto key_at = [&](uint32_t idx) -> K { return PACKED ? (K)(keys[idx] >> 32) : (K)keys[idx]; };
    for (;;) {
        const uint32_t tile = next_tile(su, &s_tile);
        if (tile >= ntiles) break;
        const uint4 td = tile_desc(g, tile / subs);
        const uint32_t lo = (tile % subs) * (kBlock * ITEMS);
        const uint32_t len = td.z > lo ? min((uint32_t)(kBlock * ITEMS), td.z - lo) : 0u;
        const uint32_t base = td.y + lo;
        const uint32_t seg_start = g.single ? 0u : g.off[td.x];
        const uint32_t end = seg_end(g, td.x);
        const uint64_t chunk = (uint64_t)td.x << lbits;
        const uint32_t w0 = wave * kWave * ITEMS;  // this wave's first element in the piece
        K k[ITEMS];
        uint32_t v[ITEMS], f[ITEMS], e[ITEMS], h[ITEMS], eh[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t li = w0 + i * kWave + lane;
            k[i] = kNone;
            v[i] = 0;
            if (li < len) {
                if constexpr (PACKED) {
                    const uint64_t w = keys[base + li];
                    k[i] = (K)(w >> 32);
                    v[i] = (uint32_t)w;
                } else {
                    k[i] = keys[base + li];
                    v[i] = vals[base + li];
                }
            }
        }
        // keys just outside this wave's elements, loaded together with them:
        // the one before its first element, the one after its last (the next
        // wave's first, or the first after the piece), and for wave 0 the 64
        // records before the piece (a brick run reaching into it)
        const uint32_t wend = min(len, w0 + (uint32_t)(kWave * ITEMS));
        const K wave_prev = w0 < len && base + w0 > seg_start ? key_at(base + w0 - 1) : kNone;
        const K wave_next = w0 < len && base + wend < end ? key_at(base + wend) : kNone;
        K back_k = kNone;
        uint32_t back_v = 0;
        if (wave == 0 && len > 0 && (int64_t)base - 1 - (int64_t)lane >= (int64_t)seg_start) {
            const uint32_t j = base - 1 - lane;
            if constexpr (PACKED) {
                const uint64_t w = keys[j];
                back_k = (K)(w >> 32);
                back_v = (uint32_t)w;
            } else {
                back_k = keys[j];
                back_v = vals[j];
            }
        }
        // kept flags (next key from the next lane, lane 63: lane 0 of the next item)
        uint64_t run_start_mask[ITEMS], kept_mask[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t li = w0 + i * kWave + lane;
            const uint32_t idx = base + li;
            const K from_lane = dppx<kDppWaveShl1>(k[i]);
            const K from_next = lane_read(k[i + 1 < ITEMS ? i + 1 : i], 0);
            const K prev_lane = dppx<kDppWaveShr1>(k[i]);
            const K prev_item = lane_read(k[i > 0 ? i - 1 : 0], kWave - 1);
            f[i] = 0;
            bool rs = false;
            if (li < len) {
                bool last = idx + 1 == end;
                if (!last) {
                    // the last element of the wave's share takes the key after it
                    const K nk = li + 1 == wend ? wave_next : lane < kWave - 1 ? from_lane : from_next;
                    last = nk != k[i];
                }
                f[i] = last && (k[i] & invalid) == 0 && v[i] != 0;
                const K pk = lane > 0 ? prev_lane : (i > 0 ? prev_item : wave_prev);
                // a new brick run starts here (or the segment starts here)
                rs = idx == seg_start || (pk >> 6) != (k[i] >> 6);
            }
            run_start_mask[i] = __ballot(rs);
            kept_mask[i] = __ballot(f[i] != 0);
        }
        // wave summary: any run start, carry out when it has one, any kept
        {
            bool any_start = false, any_kept = false, out = false;
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                const uint64_t S = run_start_mask[i], F = kept_mask[i];
                any_kept |= F != 0;
                if (S) {
                    const int r = 63 - __clzll((long long)S);  // last run start of the row
                    out = (F >> r) != 0;
                    any_start = true;
                } else if (any_start) {
                    out = out || F != 0;
                }
            }
            if (lane == 0) {
                s_any_start[wave] = any_start;
                s_any_kept[wave] = any_kept;
                s_out[wave] = out;
            }
        }
        // carry into the piece: the run reaching into element 0 from before
        if (wave == 0) {
            bool carry = false;
            const K k0 = lane_read(k[0], 0);
            // (out-of-range records never head a brick: no look through them)
            // (wave 0's wave_prev is the record before the piece)
            if (len > 0 && (k0 & invalid) == 0 && base > seg_start && (wave_prev >> 6) == (k0 >> 6)) {
                // records base-1, base-2, ... of the same brick, 64 at a time
                // (the first window came with the piece); the record after
                // each one decides whether it is the last of its key
                int64_t top = (int64_t)base - 1;
                K after = k0;  // key of the record just after the window
                K kj = back_k;
                uint32_t vj = back_v;
                for (;;) {
                    const int64_t j = top - (int64_t)lane;
                    const bool in = j >= (int64_t)seg_start;
                    // key of record j + 1 (the lane move runs in every lane:
                    // one from an inactive lane yields no defined value)
                    const K up = dppx<kDppWaveShr1>(kj);
                    const K nk = lane == 0 ? after : up;
                    const bool same = in && (kj >> 6) == (k0 >> 6);
                    const bool kept = same && nk != kj && (kj & invalid) == 0 && vj != 0;
                    if (__ballot(kept)) {
                        carry = true;
                        break;
                    }
                    const uint64_t sm = __ballot(same);
                    if (sm != ~0ull) break;  // the run starts inside this window
                    after = lane_read(kj, kWave - 1);
                    top -= kWave;
                    const int64_t jn = top - (int64_t)lane;
                    kj = kNone;
                    vj = 0;
                    if (jn >= (int64_t)seg_start) {
                        if constexpr (PACKED) {
                            const uint64_t w = keys[jn];
                            kj = (K)(w >> 32);
                            vj = (uint32_t)w;
                        } else {
                            kj = keys[jn];
                            vj = vals[jn];
                        }
                    }
                }
            }
            if (lane == 0) s_carry0 = carry;
        }
        __syncthreads();
        // this wave's carry in, then head flags row by row
        bool carry = s_carry0;
        for (uint32_t w = 0; w < wave; w++) carry = s_any_start[w] ? s_out[w] : (carry || s_any_kept[w]);
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint64_t S = run_start_mask[i], F = kept_mask[i];
            // kept records before this lane in its run, inside the row
            const uint64_t starts_upto = S & (lt | (1ull << lane));
            const int r = starts_upto ? 63 - __clzll((long long)starts_upto) : -1;
            const uint64_t span = r >= 0 ? lt & ~((1ull << r) - 1) : lt;
            const bool before = (F & span) != 0 || (r < 0 && carry);
            h[i] = f[i] && !before;
            // carry out of the row
            if (S) {
                const int rl = 63 - __clzll((long long)S);
                carry = (F >> rl) != 0;
            } else {
                carry = carry || F != 0;
            }
        }
        const uint32_t tot_u = tile_scan<ITEMS>(f, e, s_wave);
        const uint32_t tot_b = tile_scan<ITEMS>(h, eh, s_wave2);
        // both look-backs at once: wave 0 kept records, wave 1 brick heads
        if (wave == 0) {
            const uint64_t p = lookback_wave(su.status, tile, tot_u, su.epoch, &su.counters[C_ERR]);
            if (lane == 0) s_pre[0] = p;
        } else if (wave == 1) {
            const uint64_t p = lookback_wave(status_heads, tile, tot_b, su.epoch, &su.counters[C_ERR]);
            if (lane == 0) s_pre[1] = p;
        }
        __syncthreads();
        const uint64_t pre_u = s_pre[0], pre_b = s_pre[1];
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            // brick index of a kept record: heads up to it, minus one
            const uint64_t bi = pre_b + eh[i] + h[i] - 1;
            bool ok = f[i] != 0;
            if (ok) {
                const uint64_t pos = pre_u + e[i];
                if (pos >= g.n || bi >= brick_cap) {  // only after a failed look-back
                    atomicOr(&su.counters[C_ERR], 16u);
                    ok = false;
                } else {
                    ukey[pos] = chunk | k[i];
                    uval[pos] = v[i];
                    if (h[i]) brick_key[bi] = (chunk | k[i]) >> 6;
                    if (first_voxel && e[i] == 0) atomicMin(&first_voxel[td.x], (uint32_t)pos);
                }
            }
            // OR of the row's mask bits per brick, merged at each brick's last
            // kept lane: a prefix OR over lanes of the same brick (bricks are
            // contiguous in lane order; a lane of another brick, or a record
            // that is not kept, contributes nothing), steps inside rows of 16
            // and then across rows. Brick indices are below 2^32 - 1.
            uint64_t m = ok ? 1ull << (k[i] & 63) : 0;
            const uint32_t bk = ok ? (uint32_t)bi : 0xFFFFFFFFu;
            // a lane without a source reads ~bk: never equal, nothing merged
#define PYR_SEG_OR_STEP(CTRL, ROWS)                                  \
    {                                                                \
        const uint64_t y = dpp64<CTRL, ROWS>(m);                     \
        const uint32_t yb = dpp32<CTRL, ROWS>(bk, ~bk);              \
        if (yb == bk) m |= y;                                        \
    }
            PYR_SEG_OR_STEP(kDppRowShr1, 0xf)
            PYR_SEG_OR_STEP(kDppRowShr2, 0xf)
            PYR_SEG_OR_STEP(kDppRowShr4, 0xf)
            PYR_SEG_OR_STEP(kDppRowShr8, 0xf)
            PYR_SEG_OR_STEP(kDppRowBcast15, 0xa)
            PYR_SEG_OR_STEP(kDppRowBcast31, 0xc)
#undef PYR_SEG_OR_STEP
            const uint32_t nbk = dpp32<kDppWaveShl1>(bk, ~bk);
            const bool run_end = lane == kWave - 1 || nbk != bk;
            if (ok && run_end) atomicOr((unsigned long long*)&masks[bi], (unsigned long long)m);
        }
        if (tile == ntiles - 1 && threadIdx.x == 0) {
            su.counters[C_U] = (uint32_t)(pre_u + tot_u);
            su.counters[C_B] = (uint32_t)(pre_b + tot_b);
            su.counters[C_NCH + 2] = (uint32_t)(pre_b + tot_b);
        }
    }
}

// ---------------------------------------------------------------------------
// bucket finish: partial sort + one key table per bucket
// ---------------------------------------------------------------------------
// The radix passes sort only the key bits at and above `lb` (the bucket of a
// key is its segment and key >> lb; a segment of a batch has
// 2^(max(local bits of its chunk, lb) - lb) buckets, a single build 2^(local
// bits - lb)); inside a bucket the records keep their input order. The
// last record of a key is then the one with the highest index in its bucket,
// which a table of 2^lb slots in LDS finds with one atomic max per record:
// the same unique voxels (last record wins, attribute 0 deletes, out-of-range
// keys dropped) and bricks as the full sort and the unique scan
// (gbuild.zig:366-428), in the same order, without the low passes.
constexpr uint32_t kBucketBits = 14;
constexpr uint32_t kBucketSlots = 1u << kBucketBits;
constexpr uint32_t kBucketBlock = 512;
constexpr uint32_t kBucketWaves = kBucketBlock / kWave;
constexpr int kBucketLoads = 8;   // record loads per thread in flight
constexpr int kBucketGroup = 16;  // bricks per wave whose winners are fetched together
constexpr int kBucketRegs = kBucketSlots / kBucketBlock;  // records per thread of a bucket held in registers

// first_bucket[c] = buckets of the segments before c, counters[C_BKT] = all
// of them (one block). Offsets that k_seg_tiles rejected leave no bucket; a
// chunk size above log2_size counts as log2_size and raises error bit 128.
__global__ void __launch_bounds__(kBlock) k_bucket_plan(const uint8_t* chunk_log2, uint32_t count, uint32_t log2_size,
                                                         uint32_t lb, uint32_t* first_bucket, uint32_t* counters) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < count; b += kBlock) {
        const uint32_t c = b + threadIdx.x;
        uint32_t nb = 0;
        if (c < count) {
            uint32_t l2 = chunk_log2 ? chunk_log2[c] : log2_size;
            if (l2 > log2_size) {
                atomicOr(&counters[C_ERR], 128u);
                l2 = log2_size;
            }
            const uint32_t lc = local_bits(l2);
            nb = 1u << ((lc > lb ? lc : lb) - lb);
        }
        uint32_t v[1] = {nb}, e[1];
        const uint32_t sum = tile_scan<1>(v, e, s_wave);
        if (c < count) first_bucket[c] = carry + e[0];
        carry += sum;
        __syncthreads();  // s_wave is reused by the next round
    }
    if (threadIdx.x == 0) {
        first_bucket[count] = carry;
        counters[C_BKT] = counters[C_TILES] ? carry : 0u;
    }
}

// Records [start[t], end[t]) of bucket t and its segment bseg[t] (batches):
// the records of the segment whose keys lie in [j << lb, (j + 1) << lb) for
// the bucket's index j in its segment. Out-of-range keys sort after every
// bucket of their segment.
template <bool PACKED>
__global__ void __launch_bounds__(kBlock) k_bucket_bounds(const uint32_t* counters, RecordBufs b, Segs g,
                                                           const uint32_t* first_bucket, uint32_t lb,
                                                           uint32_t buckets, uint32_t* start, uint32_t* end,
                                                           uint32_t* bseg) {
    const uint64_t* keys = b.keys[counters[C_FINAL]];
    const uint32_t total = g.single ? buckets : counters[C_BKT];
    auto lower = [&](uint32_t lo, uint32_t hi, uint64_t target) {
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            const uint64_t k = PACKED ? keys[mid] >> 32 : keys[mid];
            if (k < target) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        uint32_t c = 0, j = t;
        if (!g.single) {  // the segment: last c with first_bucket[c] <= t
            uint32_t lo = 0, hi = g.count;
            while (hi - lo > 1) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (first_bucket[mid] <= t) lo = mid;
                else hi = mid;
            }
            c = lo;
            j = t - first_bucket[c];
            bseg[t] = c;
        }
        const uint32_t s0 = g.single ? 0u : g.off[c], s1 = seg_end(g, c);
        const uint32_t a = lower(s0, s1, (uint64_t)j << lb);
        start[t] = a;
        end[t] = lower(a, s1, (uint64_t)(j + 1) << lb);
    }
}

// One bucket per tile, tiles handed out in order: clear the table, record
// per key the highest index + 1 of its records, put each winner's attribute
// in its slot (kept when non-zero; a bucket of at most 2^14 packed records
// stays in registers, so the winners write their attribute after a compare,
// larger ones fetch it again), count kept voxels and non-empty bricks per brick of
// 64 slots, two decoupled look-backs for the bucket's first unique position
// and brick index (waves 0 and 1), then write keys, attributes, brick keys
// and masks. A brick lies inside one bucket, so its mask is written whole.
template <bool PACKED>
__global__ void __launch_bounds__(kBucketBlock, 4)
    k_bucket_unique(ScanCtx su, uint64_t* status_heads, const uint32_t* start, const uint32_t* end,
                    const uint32_t* bseg, uint32_t buckets, RecordBufs b, uint32_t lb, uint32_t lbits,
                    uint64_t* brick_key, uint64_t* masks, uint32_t* first_voxel, uint64_t brick_cap, uint32_t n_cap) {
    constexpr uint32_t kBricksMax = kBucketSlots / 64;
    __shared__ uint32_t tab[kBucketSlots];
    __shared__ uint64_t s_mask[kBricksMax];
    __shared__ uint32_t s_bu[kBricksMax], s_bh[kBricksMax];
    __shared__ uint32_t s_tile, s_wu[kBucketWaves], s_wh[kBucketWaves];
    __shared__ uint64_t s_pre[2];
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    const uint64_t lt = (1ull << lane) - 1;
    const uint32_t slots = 1u << lb, nbr = slots >> 6;
    // slot of a key in the table: the bit index inside its brick XOR the
    // brick's low six bits, so that keys of neighbouring bricks (records in
    // input order, e.g. along a column) fall into different LDS banks; a
    // brick keeps its 64 words
    auto phys = [](uint32_t k) -> uint32_t { return k ^ ((k >> 6) & 63u); };
    const uint32_t fin = su.counters[C_FINAL];
    const uint64_t* keys = b.keys[fin];
    const uint32_t* vals = b.vals[fin];
    uint64_t* ukey = b.keys[fin ^ 1];
    uint32_t* uval = PACKED ? b.vals[0] : b.vals[fin ^ 1];
    const uint32_t total = bseg ? su.counters[C_BKT] : buckets;
    for (;;) {
        const uint32_t bk = next_tile(su, &s_tile);
        if (bk >= total) break;
        const uint32_t s = start[bk], e = end[bk];
        uint32_t tot_u = 0, tot_b = 0;
        if (s < e) {  // (uniform in the block)
            for (uint32_t i = threadIdx.x; i < slots / 4; i += kBucketBlock)
                reinterpret_cast<uint4*>(tab)[i] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
            if (PACKED && e - s <= kBucketSlots) {
                // the whole bucket in registers: winners take their attribute
                // straight from them (no second read of the records)
                uint32_t sl[kBucketRegs], at[kBucketRegs];  // slot (~0: no record) and attribute
#pragma unroll
                for (int u = 0; u < kBucketRegs; u++) {
                    const uint32_t j = s + u * kBucketBlock + threadIdx.x;
                    const uint64_t w = j < e ? __builtin_nontemporal_load(&keys[j]) : ~0ull;
                    sl[u] = j < e ? phys((uint32_t)(w >> 32) & (slots - 1)) : ~0u;
                    at[u] = (uint32_t)w;
                }
#pragma unroll
                for (int u = 0; u < kBucketRegs; u++)
                    if (sl[u] != ~0u)
                        __hip_atomic_fetch_max(&tab[sl[u]], u * kBucketBlock + threadIdx.x + 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                __syncthreads();
                uint32_t win = 0;
#pragma unroll
                for (int u = 0; u < kBucketRegs; u++)
                    if (sl[u] != ~0u && tab[sl[u]] == u * kBucketBlock + threadIdx.x + 1) win |= 1u << u;
                __syncthreads();  // every comparison is done before a winner overwrites its slot
#pragma unroll
                for (int u = 0; u < kBucketRegs; u++)
                    if ((win >> u) & 1) tab[sl[u]] = at[u];
            } else {
                for (uint32_t j0 = s; j0 < e; j0 += kBucketBlock * kBucketLoads) {
                    uint64_t kk[kBucketLoads];
#pragma unroll
                    for (int u = 0; u < kBucketLoads; u++) {
                        const uint32_t j = j0 + u * kBucketBlock + threadIdx.x;
                        kk[u] = j < e ? keys[j] : 0;
                    }
#pragma unroll
                    for (int u = 0; u < kBucketLoads; u++) {
                        const uint32_t j = j0 + u * kBucketBlock + threadIdx.x;
                        if (j < e) {
                            const uint64_t key = PACKED ? kk[u] >> 32 : kk[u];
                            // (records of another bucket cannot be here; out-of-range
                            // keys lie past start[buckets])
                            __hip_atomic_fetch_max(&tab[phys((uint32_t)key & (slots - 1))], j - s + 1,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
                __syncthreads();
                // winners' attributes into the table
                for (uint32_t g0 = wave * kBucketGroup; g0 < nbr; g0 += kBucketWaves * kBucketGroup) {
                    uint32_t idx[kBucketGroup], v[kBucketGroup];
#pragma unroll
                    // (word by word: the slot layout does not matter here)
                    for (int q = 0; q < kBucketGroup; q++) idx[q] = g0 + q < nbr ? tab[(g0 + q) * 64 + lane] : 0u;
#pragma unroll
                    for (int q = 0; q < kBucketGroup; q++) {
                        v[q] = 0;
                        if (idx[q]) v[q] = PACKED ? (uint32_t)keys[s + idx[q] - 1] : vals[s + idx[q] - 1];
                    }
#pragma unroll
                    for (int q = 0; q < kBucketGroup; q++)
                        if (g0 + q < nbr) tab[(g0 + q) * 64 + lane] = v[q];
                }
            }
            __syncthreads();
            // brick masks of kept voxels (attribute non-zero)
            for (uint32_t br = wave; br < nbr; br += kBucketWaves) {
                const uint64_t m = __ballot(tab[phys(br * 64 + lane)] != 0);
                if (lane == 0) s_mask[br] = m;
            }
            __syncthreads();
            // exclusive prefixes of kept voxels and non-empty bricks over the bricks
            uint32_t cu = 0, ch = 0;
            if (threadIdx.x < nbr) {
                const uint64_t m = s_mask[threadIdx.x];
                cu = (uint32_t)__popcll(m);
                ch = m != 0;
            }
            const uint32_t iu = wave_inclusive_scan(cu), ih = wave_inclusive_scan(ch);
            if (lane == kWave - 1) {
                s_wu[wave] = iu;
                s_wh[wave] = ih;
            }
            __syncthreads();
            uint32_t ou = 0, oh = 0;
            for (uint32_t w = 0; w < kBucketWaves; w++) {
                if (w < wave) {
                    ou += s_wu[w];
                    oh += s_wh[w];
                }
                tot_u += s_wu[w];
                tot_b += s_wh[w];
            }
            if (threadIdx.x < nbr) {
                s_bu[threadIdx.x] = ou + iu - cu;
                s_bh[threadIdx.x] = oh + ih - ch;
            }
        }
        if (wave == 0) {
            const uint64_t p = lookback_wave(su.status, bk, tot_u, su.epoch, &su.counters[C_ERR]);
            if (lane == 0) s_pre[0] = p;
        } else if (wave == 1) {
            const uint64_t p = lookback_wave(status_heads, bk, tot_b, su.epoch, &su.counters[C_ERR]);
            if (lane == 0) s_pre[1] = p;
        }
        __syncthreads();
        const uint64_t pre_u = s_pre[0], pre_b = s_pre[1];
        if (tot_u) {
            // the segment's key prefix and the bucket's key bits (from its first record)
            const uint32_t c = bseg ? bseg[bk] : 0u;
            const uint64_t k0 = PACKED ? keys[s] >> 32 : keys[s];
            const uint64_t base_key = ((uint64_t)c << lbits) | ((k0 >> lb) << lb);
            if (first_voxel && threadIdx.x == 0) atomicMin(&first_voxel[c], (uint32_t)pre_u);
            for (uint32_t br = wave; br < nbr; br += kBucketWaves) {
                const uint64_t m = s_mask[br];
                if (!m) continue;  // (uniform in the wave)
                const uint64_t bi = pre_b + s_bh[br], pos = pre_u + s_bu[br] + __popcll(m & lt);
                const uint64_t key0 = base_key | ((uint64_t)br << 6);
                if (bi >= brick_cap || pre_u + s_bu[br] + __popcll(m) > n_cap) {  // only after a failed look-back
                    if (lane == 0) atomicOr(&su.counters[C_ERR], 16u);
                    continue;
                }
                if (lane == 0) {
                    brick_key[bi] = key0 >> 6;
                    masks[bi] = m;
                }
                if ((m >> lane) & 1) {
                    ukey[pos] = key0 | lane;
                    uval[pos] = tab[phys(br * 64 + lane)];
                }
            }
        }
        if (bk == total - 1 && threadIdx.x == 0) {
            su.counters[C_U] = (uint32_t)(pre_u + tot_u);
            su.counters[C_B] = (uint32_t)(pre_b + tot_b);
            su.counters[C_NCH + 2] = (uint32_t)(pre_b + tot_b);
        }
    }
}

// clear the first hash_cap_for(count, load) slots of a table (keys 0, owner ~0)
__global__ void __launch_bounds__(kBlock) k_hash_clear(const uint32_t* counters, uint32_t count_slot, void* table,
                                                        uint32_t key_bytes, uint32_t* owner, uint64_t max_cap,
                                                        uint32_t load_pct) {
    const uint64_t want = hash_cap_for(counters[count_slot], load_pct);
    const uint64_t cap = want < max_cap ? want : max_cap;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cap; i += stride) {
        if (key_bytes == 8) ((uint64_t*)table)[i] = 0;
        else ((uint32_t*)table)[i] = 0;
        owner[i] = 0xFFFFFFFFu;
    }
}

// Owner = smallest index (gbuild.zig:450-454). The atomic is skipped when the
// slot already holds an index <= i: a stale read can only be larger than the
// current value, so skipping never loses a smaller index.
__device__ __forceinline__ void owner_min(uint32_t* owner, uint64_t h, uint32_t i) {
    if (load_agent(&owner[h]) > i) atomicMin(&owner[h], i);
}

// leaf_insert + owner (gbuild.zig:436-454): CAS of the mask (0 = empty slot),
// then the smallest brick index with that mask becomes its owner. Solid
// terrain makes one mask (the full brick) dominate; lanes of a wave holding
// the same mask as the wave's first lane reuse its slot, and slots are read
// before any CAS, so a hot mask costs loads instead of serialised atomics.
// The wave also keeps the mask and slot of its previous row: a lane holding
// that mask takes the slot without touching the table. Its brick index is
// larger than the one of the earlier lane that competed for the owner, so
// skipping the owner update loses nothing. (The device-coherent reads of one
// hot slot by every wave of the grid serialise at one memory channel.)
// Workgroups of the hash inserts. Every wave's first row makes the hot key's
// device-scope CAS and owner atomic (later rows reuse the slot), and atomics
// on one address are serialised, so fewer waves, each with more rows: leaf
// insert of the 1024^3 grid 0.32 ms at 4096 workgroups, 0.31 at 2048, 0.22 at
// 1024, 0.17 at 512 and 0.16 at 256; of the 1024-chunk batch 0.43 / 0.30 /
// 0.25 / 0.28 ms at 2048 / 1024 / 512 / 256.
constexpr uint32_t kInsertBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_leaf_insert(uint32_t* counters, const uint64_t* masks, uint64_t* table,
                                                         uint32_t* owner, uint32_t* slot, uint32_t* ch_count,
                                                         uint64_t max_cap, uint32_t brick_cap, uint32_t load_pct) {
    const uint32_t n = min(counters[C_B], brick_cap);
    const uint64_t want = hash_cap_for(n, load_pct);
    const uint64_t cap = want < max_cap ? want : max_cap, m = cap - 1;
    const uint32_t stride = gridDim.x * kBlock;
    uint64_t c_mask = 0, c_slot = 0;  // previous row's first mask and slot (masks are non-zero)
    uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    uint64_t next = b < n ? masks[b] : 0;
    for (; b < n; b += stride) {
        const uint64_t mask = next;
        if (b + stride < n) next = masks[b + stride];  // the next row's load overlaps this row's probes
        const uint64_t m0 = first_lane64(mask);
        const uint32_t b0 = first_lane32(b);
        const bool follower = mask == m0 && b != b0;
        const bool cached = mask == c_mask;
        uint64_t h = cached ? c_slot : mix64(mask) & m;
        if (!follower && !cached) {
            uint64_t probes = 0;
            for (;;) {
                uint64_t cur = load_agent(&table[h]);
                if (cur == 0) cur = atomicCAS((unsigned long long*)&table[h], 0ull, (unsigned long long)mask);
                if (cur == 0 || cur == mask) break;
                h = (h + 1) & m;
                if (++probes > cap) {
                    atomicOr(&counters[C_ERR], 4u);
                    break;
                }
            }
            owner_min(owner, h, b);
        }
        const uint64_t h0 = first_lane64(h);  // the slot of the wave's first active lane
        if (follower) h = h0;  // b0 < b already competes for the owner
        slot[b] = (uint32_t)h;
        ch_count[b] = (uint32_t)__popcll(mask);
        c_mask = m0;
        c_slot = h0;
    }
}

// leaf_first/leaf_write (gbuild.zig:456-469): owners get leaf indices in brick
// order; every brick records its owner
__global__ void __launch_bounds__(kBlock) k_leaf_compact(ScanCtx sc, const uint64_t* masks, const uint32_t* owner,
                                                          const uint32_t* slot, uint64_t* leaves_out,
                                                          uint32_t* ch_rep, uint32_t* ch_pos, uint32_t brick_cap) {
    __shared__ uint32_t s_tile, s_wave[4];
    __shared__ uint64_t s_b;
    const uint32_t n = min(sc.counters[C_B], brick_cap);
    const uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
    for (;;) {
        const uint32_t tile = next_tile(sc, &s_tile);
        if (tile >= ntiles) break;
        uint32_t f[kScanItems], e[kScanItems], o[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            uint64_t idx = scan_index(tile, i);
            f[i] = 0;
            if (idx < n) {
                o[i] = owner[slot[idx]];
                f[i] = o[i] == (uint32_t)idx;
            }
        }
        uint32_t total = tile_scan<kScanItems>(f, e, s_wave);
        uint64_t pre = tile_prefix(sc, tile, total, &s_b);
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            uint64_t idx = scan_index(tile, i);
            if (idx < n) {
                ch_rep[idx] = o[i];
                if (f[i]) {
                    const uint64_t pos = pre + e[i];
                    if (pos >= n) {  // only after a failed look-back
                        atomicOr(&sc.counters[C_ERR], 16u);
                        continue;
                    }
                    leaves_out[pos] = masks[idx];
                    ch_pos[idx] = (uint32_t)pos;
                }
            }
        }
        if (tile == ntiles - 1 && threadIdx.x == 0) sc.counters[C_LEAVES] = (uint32_t)(pre + total);
    }
}

// One element set of a level: keys, the owner (representative) of each
// element, the final ref of each owner, and voxel counts.
struct LevelArrays {
    uint64_t* key;
    uint32_t* rep;
    uint32_t* pos;
    uint32_t* count;
};

struct LevelArgs {
    uint32_t lv;              // children have edge 2^lv
    uint32_t top_shift;       // 3 * (log2_size - lv): key >> top_shift = chunk (children level)
    const uint8_t* chunk_log2;
    uint32_t* roots;          // chunk roots (batches) or null
    LevelArrays ch, par;
    uint32_t* node_tmp;       // [parents][10]
    uint32_t* node_len;
    uint32_t* table;          // tuple hash: index + 1, 0 = empty
    uint32_t* owner;
    uint32_t* slot;
    // capacities of the arrays above (guards against corrupt counts)
    uint32_t ch_cap, par_cap;
    uint64_t table_cap, word_cap;
    uint32_t chunk_count;
    uint32_t load_pct;        // maximum load of the tuple hash in percent
};

// Child idx of a level (key k) heads a parent when the parent key k >> 3
// changes there (gbuild.zig:475-480); a chunk whose edge is 2^lv ends at this
// level instead: its only element is the chunk's root, recorded in roots.
__device__ __forceinline__ uint32_t level_head(const LevelArgs& a, uint64_t idx, uint64_t k) {
    if (a.chunk_log2) {
        const uint32_t c = (uint32_t)(k >> a.top_shift);
        if (c < a.chunk_count && a.chunk_log2[c] == a.lv) {
            a.roots[c] = a.ch.pos[a.ch.rep[idx]];
            return 0;
        }
    }
    return idx == 0 || (a.ch.key[idx - 1] >> 3) != (k >> 3);
}

// node_build (gbuild.zig:482-520): the head at idx gathers its (at most 8)
// children into parent p's tuple [mask, count, child positions]; their keys,
// then reps and counts, then positions are each loaded together
__device__ __forceinline__ void level_tuple(const LevelArgs& a, uint64_t idx, uint32_t n, uint64_t k, uint32_t p) {
    const uint64_t pk = k >> 3;
    uint32_t w[kNodeStride];
    uint32_t mask = 0, count = 0, len = 2;
    uint64_t ck[8];
    uint32_t nkids = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) ck[q] = idx + q < n ? a.ch.key[idx + q] : ~0ull;
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (nkids == (uint32_t)q && (ck[q] >> 3) == pk) nkids++;
    uint32_t rep[8], cnt[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        rep[q] = (uint32_t)q < nkids ? a.ch.rep[idx + q] : 0u;
        cnt[q] = (uint32_t)q < nkids ? a.ch.count[idx + q] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if ((uint32_t)q < nkids) {
            mask |= 1u << (uint32_t)(ck[q] & 7);
            w[len++] = a.ch.pos[rep[q]];
            count = sat_add(count, cnt[q]);
        }
    }
    w[0] = mask;
    w[1] = count;
    uint32_t* dst = a.node_tmp + (uint64_t)p * kNodeStride;
    for (uint32_t q = 0; q < len; q++) dst[q] = w[q];
    a.node_len[p] = len;
    a.par.key[p] = pk;
    a.par.count[p] = count;
}

// node_insert/node_owner (gbuild.zig:527-554) of parent p's tuple w (len
// words): the level's hash holds index + 1; a collision compares the stored
// tuple word by word; the owner of a slot is the smallest index. Returns the
// slot.
__device__ __forceinline__ uint64_t level_insert_one(const LevelArgs& a, uint32_t p, const uint32_t* w, uint32_t len,
                                                     uint64_t cap, uint32_t* err) {
    const uint64_t hm = cap - 1;
    uint64_t hv = len;
    for (uint32_t q = 0; q < len; q++) hv = mix64(hv ^ w[q]);
    uint64_t h = hv & hm;
    uint64_t probes = 0;
    for (;;) {
        uint32_t cur = load_agent(&a.table[h]);
        if (cur == 0) {
            cur = atomicCAS(&a.table[h], 0u, p + 1);
            if (cur == 0) break;
        }
        const uint32_t other = cur - 1;
        if (a.node_len[other] == len) {
            const uint32_t* ow = a.node_tmp + (uint64_t)other * kNodeStride;
            bool same = true;
            for (uint32_t q = 0; q < len; q++) same &= ow[q] == w[q];
            if (same) break;
        }
        h = (h + 1) & hm;
        if (++probes > cap) {
            atomicOr(err, 8u);
            break;
        }
    }
    owner_min(a.owner, h, p);
    return h;
}

// node_heads/node_build/node_insert/node_owner (gbuild.zig:475-550) in one
// pass: parents are numbered by a tile scan of the head flags, each head
// gathers its <= 8 children, and the tuple is inserted into the level's hash
__global__ void __launch_bounds__(kBlock) k_level_build(ScanCtx sc, LevelArgs a) {
    __shared__ uint32_t s_tile, s_wave[4];
    __shared__ uint64_t s_b;
    const uint32_t n = min(sc.counters[C_NCH + a.lv], a.ch_cap);
    const uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
    for (;;) {
        const uint32_t tile = next_tile(sc, &s_tile);
        if (tile >= ntiles) break;
        uint32_t f[kScanItems], e[kScanItems];
        uint64_t k[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            uint64_t idx = scan_index(tile, i);
            f[i] = 0;
            if (idx < n) {
                k[i] = a.ch.key[idx];
                f[i] = level_head(a, idx, k[i]);
            }
        }
        uint32_t total = tile_scan<kScanItems>(f, e, s_wave);
        uint64_t pre = tile_prefix(sc, tile, total, &s_b);
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            if (!f[i]) continue;
            const uint64_t idx = scan_index(tile, i);
            if (pre + e[i] >= a.par_cap) {  // only after a failed look-back
                atomicOr(&sc.counters[C_ERR], 16u);
                continue;
            }
            level_tuple(a, idx, n, k[i], (uint32_t)(pre + e[i]));
        }
        if (tile == ntiles - 1 && threadIdx.x == 0) sc.counters[C_NCH + a.lv + 1] = (uint32_t)(pre + total);
    }
}

// node_insert/node_owner (gbuild.zig:527-554): the tuples written by
// k_level_build (complete at this launch) are inserted into the level's hash
// as index + 1; a collision compares the stored tuple word by word. Hot tuples
// (identical full nodes of solid regions) are handled like hot leaf masks,
// the previous row's first tuple and slot included.
__global__ void __launch_bounds__(kBlock) k_level_insert(uint32_t* counters, LevelArgs a) {
    const uint32_t np = min(counters[C_NCH + a.lv + 1], a.par_cap);
    const uint32_t nch = min(counters[C_NCH + a.lv], a.ch_cap);
    const uint64_t want = hash_cap_for(nch, a.load_pct);
    const uint64_t cap = want < a.table_cap ? want : a.table_cap;
    const uint32_t stride = gridDim.x * kBlock;
    uint32_t c_len = ~0u, c_w[kNodeStride] = {};  // previous row's first tuple (lengths are <= 8)
    uint64_t c_slot = 0;
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < np; p += stride) {
        const uint32_t len = a.node_len[p];
        const uint32_t* src = a.node_tmp + (uint64_t)p * kNodeStride;
        uint32_t w[kNodeStride];
#pragma unroll
        for (int q = 0; q < kNodeStride; q++) w[q] = (uint32_t)q < len ? src[q] : 0u;
        // same tuple as the wave's first lane?
        const uint32_t p0 = first_lane32(p);
        bool same0 = first_lane32(len) == len;
#pragma unroll
        for (int q = 0; q < kNodeStride; q++) same0 &= first_lane32(w[q]) == w[q];
        const bool follower = same0 && p != p0;
        bool cached = len == c_len;
#pragma unroll
        for (int q = 0; q < kNodeStride; q++) cached &= c_w[q] == w[q];
        uint64_t h = c_slot;
        if (!follower && !cached) h = level_insert_one(a, p, w, len, cap, &counters[C_ERR]);
        const uint64_t h0 = first_lane64(h);  // the slot of the wave's first active lane
        if (follower) h = h0;
        a.slot[p] = (uint32_t)h;
        c_len = first_lane32(len);
#pragma unroll
        for (int q = 0; q < kNodeStride; q++) c_w[q] = first_lane32(w[q]);
        c_slot = h0;
    }
}

// node_first/node_write (gbuild.zig:552-571): owners' words are appended at
// the level cursor in parent order; every parent records its owner
__global__ void __launch_bounds__(kBlock) k_level_compact(ScanCtx sc, LevelArgs a, uint32_t* nodes_out) {
    __shared__ uint32_t s_tile, s_wave[4];
    __shared__ uint64_t s_b;
    const uint32_t n = min(sc.counters[C_NCH + a.lv + 1], a.par_cap);
    const uint32_t cur = sc.counters[C_CUR + a.lv];
    const uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) sc.counters[C_CUR + a.lv + 1] = cur;
    for (;;) {
        const uint32_t tile = next_tile(sc, &s_tile);
        if (tile >= ntiles) break;
        uint32_t f[kScanItems], e[kScanItems], o[kScanItems];
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            uint64_t idx = scan_index(tile, i);
            f[i] = 0;
            if (idx < n) {
                o[i] = a.owner[a.slot[idx]];
                f[i] = o[i] == (uint32_t)idx ? a.node_len[idx] : 0;
            }
        }
        uint32_t total = tile_scan<kScanItems>(f, e, s_wave);
        uint64_t pre = tile_prefix(sc, tile, total, &s_b);
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            uint64_t idx = scan_index(tile, i);
            if (idx >= n) continue;
            a.par.rep[idx] = o[i];
            if (f[i]) {
                const uint32_t at = cur + (uint32_t)(pre + e[i]);
                if ((uint64_t)at + f[i] > a.word_cap) {  // only after a failed look-back
                    atomicOr(&sc.counters[C_ERR], 16u);
                    continue;
                }
                const uint32_t* src = a.node_tmp + idx * kNodeStride;
                for (uint32_t q = 0; q < f[i]; q++) nodes_out[at + q] = src[q];
                a.par.pos[idx] = at;
            }
        }
        if (tile == ntiles - 1 && threadIdx.x == 0) sc.counters[C_CUR + a.lv + 1] = cur + (uint32_t)(pre + total);
    }
}

// Every node level of a small build in one workgroup (1024 threads) and one
// launch: per level the hash is cleared, heads are numbered by a block scan
// carried from round to round, tuples are gathered and inserted, and the
// owners' words are appended at the level cursor, the same arrays and
// counters as k_hash_clear, k_level_build, k_level_insert and
// k_level_compact write, with barriers instead of launches and look-backs.
// `a` describes the first level (children of edge 4 in set ch, parents in
// par); the sets swap from level to level like the launches' arguments.
constexpr uint32_t kLevelsBlock = 1024;
constexpr uint64_t kLevelsSmallBricks = 32768;

__device__ __forceinline__ uint32_t levels_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    constexpr uint32_t W = kLevelsBlock / kWave;
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, t = 0;
#pragma unroll
    for (uint32_t q = 0; q < W; q++) {
        const uint32_t x = s_wave[q];
        before += q < wave ? x : 0u;
        t += x;
    }
    __syncthreads();  // s_wave is written again by the next round
    *total = t;
    return before + inc - v;
}

__global__ void __launch_bounds__(kLevelsBlock) k_levels_small(uint32_t* counters, LevelArgs a, uint32_t log2_size,
                                                               uint32_t cap0, uint32_t cap1, uint32_t par_max,
                                                               uint32_t* nodes_out) {
    __shared__ uint32_t s_wave[kLevelsBlock / kWave];
    const uint32_t tid = threadIdx.x;
    uint32_t caps[2] = {cap0, cap1};
    int ch = 0;
    for (uint32_t lv = 2; lv < log2_size; lv++) {
        a.lv = lv;
        a.top_shift = 3 * (log2_size - lv);
        a.ch_cap = caps[ch];
        a.par_cap = min(caps[ch ^ 1], par_max);
        const uint32_t nch = min(counters[C_NCH + lv], a.ch_cap);
        // hash of this level (k_hash_clear)
        const uint64_t want = hash_cap_for(nch, a.load_pct);
        const uint64_t cap = want < a.table_cap ? want : a.table_cap;
        for (uint64_t i = tid; i < cap; i += kLevelsBlock) {
            a.table[i] = 0;
            a.owner[i] = 0xFFFFFFFFu;
        }
        // heads and tuples (k_level_build)
        uint32_t np = 0;
        for (uint32_t base = 0; base < nch; base += kLevelsBlock) {
            const uint32_t idx = base + tid;
            uint64_t k = 0;
            uint32_t f = 0;
            if (idx < nch) {
                k = a.ch.key[idx];
                f = level_head(a, idx, k);
            }
            uint32_t total;
            const uint32_t e = levels_scan(f, s_wave, &total);
            if (f) {
                if (np + e >= a.par_cap) atomicOr(&counters[C_ERR], 16u);
                else level_tuple(a, idx, nch, k, np + e);
            }
            np += total;
        }
        __syncthreads();  // tuples and the cleared hash are complete
        if (tid == 0) counters[C_NCH + lv + 1] = np;
        np = min(np, a.par_cap);
        // insert (k_level_insert)
        for (uint32_t p = tid; p < np; p += kLevelsBlock) {
            const uint32_t len = a.node_len[p];
            const uint32_t* src = a.node_tmp + (uint64_t)p * kNodeStride;
            uint32_t w[kNodeStride];
#pragma unroll
            for (int q = 0; q < kNodeStride; q++) w[q] = (uint32_t)q < len ? src[q] : 0u;
            a.slot[p] = (uint32_t)level_insert_one(a, p, w, len, cap, &counters[C_ERR]);
        }
        __syncthreads();  // every owner is final
        // owners' words at the level cursor (k_level_compact)
        const uint32_t cur = counters[C_CUR + lv];
        uint32_t at0 = 0;
        for (uint32_t base = 0; base < np; base += kLevelsBlock) {
            const uint32_t idx = base + tid;
            uint32_t f = 0, o = 0;
            if (idx < np) {
                o = a.owner[a.slot[idx]];
                f = o == idx ? a.node_len[idx] : 0u;
            }
            uint32_t total;
            const uint32_t e = levels_scan(f, s_wave, &total);
            if (idx < np) {
                a.par.rep[idx] = o;
                if (f) {
                    const uint32_t at = cur + at0 + e;
                    if ((uint64_t)at + f > a.word_cap) {
                        atomicOr(&counters[C_ERR], 16u);
                    } else {
                        const uint32_t* src = a.node_tmp + (uint64_t)idx * kNodeStride;
                        for (uint32_t q = 0; q < f; q++) nodes_out[at + q] = src[q];
                        a.par.pos[idx] = at;
                    }
                }
            }
            at0 += total;
        }
        if (tid == 0) counters[C_CUR + lv + 1] = cur + at0;
        __syncthreads();  // the next level reads this one's counters and parents
        const LevelArrays t = a.ch;
        a.ch = a.par;
        a.par = t;
        ch ^= 1;
    }
}

// roots of the top level and the empty geometry (gbuild.zig:607-624)
__global__ void __launch_bounds__(kBlock) k_finalize(uint32_t* counters, uint32_t top, LevelArrays par,
                                                      uint32_t* roots, uint32_t* nodes_out, uint32_t par_cap,
                                                      uint32_t chunk_count, uint64_t word_cap) {
    const uint32_t np = min(counters[C_NCH + top], par_cap);
    const uint32_t cur = counters[C_CUR + top];
    if (roots) {
        for (uint32_t p = threadIdx.x; p < np; p += kBlock) {
            const uint64_t c = par.key[p];
            if (c < chunk_count) roots[c] = par.pos[par.rep[p]];
            else atomicOr(&counters[C_ERR], 16u);
        }
    }
    if (threadIdx.x == 0) {
        uint32_t root = np > 0 ? par.pos[par.rep[0]] : 0;
        uint32_t words = cur;
        if (counters[C_U] == 0 && (uint64_t)cur + 2 <= word_cap) {
            nodes_out[cur] = 0;
            nodes_out[cur + 1] = 0;
            root = cur;
            words = cur + 2;
        }
        counters[C_ROOT] = root;
        counters[C_WORDS] = words;
    }
}

__device__ __forceinline__ uint32_t compact3(uint64_t v) {
    uint64_t x = v & 0x1249249249249249ull;
    x = (x | (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x | (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x | (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x | (x >> 16)) & 0x1F00000000FFFFull;
    x = (x | (x >> 32)) & 0x1FFFFFull;
    return (uint32_t)x;
}

// sorted keys (brick Morton << 6 | bit, chunk bits above lbits dropped) back to
// PyrVoxel records in chunk-local coordinates
__global__ void __launch_bounds__(kBlock) k_decode(const uint64_t* keys, const uint32_t* attrs, uint32_t count,
                                                    uint32_t lbits, uint32_t shift, int4* out) {
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const uint64_t key = keys[i] & ((1ull << lbits) - 1);
        const uint64_t brick = key >> 6;
        const uint32_t bit = (uint32_t)(key & 63);
        int4 v;
        // coordinates, moved to their cells of edge 2^shift (downsample, gbuild.zig:262-274)
        v.x = (int)(((compact3(brick) << 2) | (bit & 3)) >> shift);
        v.y = (int)(((compact3(brick >> 1) << 2) | ((bit >> 2) & 3)) >> shift);
        v.z = (int)(((compact3(brick >> 2) << 2) | (bit >> 4)) >> shift);
        v.w = (int)attrs[i];
        out[i] = v;
    }
}

// plain streaming copy of n 16-byte words: the calibration kernel for the HBM
// byte counters (bytes read = bytes written = 16n) and the copy ceiling.
// 16-byte words per lane (the widest global access), non-temporal loads and
// stores, each block copying one contiguous 256 KiB chunk with eight loads in
// flight per lane: the fastest of the copy shapes measured on the MI355X
// (5.5 TB/s against 4.6 TB/s for an 8-byte grid-stride copy)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr uint64_t kProbeChunk = (256u << 10) / 16;  // 16-byte words per block
__global__ void __launch_bounds__(kBlock) k_copy_probe(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                        uint64_t n) {
    constexpr int U = 8;
    const uint64_t lo = (uint64_t)blockIdx.x * kProbeChunk, hi = min(lo + kProbeChunk, n);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += (uint64_t)kBlock * U) {
        u32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (i + u * kBlock < hi) v[u] = __builtin_nontemporal_load(&src[i + u * kBlock]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (i + u * kBlock < hi) __builtin_nontemporal_store(v[u], &dst[i + u * kBlock]);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
namespace {

inline uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap) {
    uint64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    return (uint32_t)std::min<uint64_t>(g, cap);
}

void fill32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256 * 4, 4096)), dim3(256), 0, s, p, v, n);
}
void fill64(hipStream_t s, uint64_t* p, uint64_t v, uint64_t n) {
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256 * 4, 4096)), dim3(256), 0, s, p, v, n);
}

// Device memory of one build: from the library's cache, or (base set) carved
// out of a fixed device range in 256-byte aligned pieces, with no allocation.
struct Alloc {
    std::vector<void*> ptrs;
    hipStream_t s;
    bool failed = false;
    uint8_t* base = nullptr;
    uint64_t size = 0, used = 0;
    uint64_t requested = 0;  // bytes asked for so far
    void* raw(uint64_t bytes) {
        bytes = std::max<uint64_t>(bytes, 256);
        requested += bytes;
        if (base) {
            const uint64_t at = (used + 255) & ~255ull;
            if (at + bytes > size) {
                failed = true;
                used = at + bytes;  // keep counting: the error reports the need
                return nullptr;
            }
            used = at + bytes;
            return base + at;
        }
        void* p = nullptr;
        if (pool_alloc(&p, bytes, s) != hipSuccess) {
            failed = true;
            return nullptr;
        }
        ptrs.push_back(p);
        return p;
    }
    template <class T>
    T* get(uint64_t count) {
        return (T*)raw(count * sizeof(T));
    }
    void release() {
        for (void* p : ptrs) pool_free(p, s);
        ptrs.clear();
    }
};

// timing switch and the last build's phase times (builds may run on several
// threads at once, e.g. the build server's workers)
std::atomic<bool> g_timing{false};
std::mutex g_times_mutex;
PyrhipBuildTimes g_times{};

struct Timer {
    hipEvent_t ev[8];
    int n = 0;
    hipStream_t s;
    bool on = g_timing;
    void mark() {
        if (!on) return;
        if (hipEventCreate(&ev[n]) != hipSuccess) {  // timing is best effort: stop marking
            (void)hipGetLastError();
            on = false;
            return;
        }
        (void)hipEventRecord(ev[n], s);
        n++;
    }
    float between(int a, int b) {
        float ms = 0;
        if (on && b < n && hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0;
        }
        return ms;
    }
    ~Timer() {
        for (int i = 0; i < n; i++) (void)hipEventDestroy(ev[i]);
    }
};

template <int RB, bool PACKED, bool NARROW>
void launch_sort_rb(hipStream_t s, uint32_t items, uint32_t tiles, const SortArgs& a) {
    switch (items) {
        case 8: hipLaunchKernelGGL((k_sort_pass<RB, 8, PACKED, NARROW>), dim3(tiles), dim3(kBlock), 0, s, a); break;
        case 12: hipLaunchKernelGGL((k_sort_pass<RB, 12, PACKED, NARROW>), dim3(tiles), dim3(kBlock), 0, s, a); break;
        case 20: hipLaunchKernelGGL((k_sort_pass<RB, 20, PACKED, NARROW>), dim3(tiles), dim3(kBlock), 0, s, a); break;
        case 24: hipLaunchKernelGGL((k_sort_pass<RB, 24, PACKED, NARROW>), dim3(tiles), dim3(kBlock), 0, s, a); break;
        default: hipLaunchKernelGGL((k_sort_pass<RB, 16, PACKED, NARROW>), dim3(tiles), dim3(kBlock), 0, s, a); break;
    }
}

// packed records are sorted with 8-bit digits only (the fastest width), with
// narrow look-back words when a.status32 is set
void launch_sort(hipStream_t s, uint32_t rb, uint32_t items, bool packed, uint32_t tiles, const SortArgs& a) {
    if (packed) {
        if (a.status32) launch_sort_rb<8, true, true>(s, items, tiles, a);
        else launch_sort_rb<8, true, false>(s, items, tiles, a);
        return;
    }
    switch (rb) {
        case 8: launch_sort_rb<8, false, false>(s, items, tiles, a); break;
        case 9: launch_sort_rb<9, false, false>(s, items, tiles, a); break;
        case 10: launch_sort_rb<10, false, false>(s, items, tiles, a); break;
        default: launch_sort_rb<11, false, false>(s, items, tiles, a); break;
    }
}

template <int ITEMS>
void launch_unique_bricks_t(hipStream_t s, uint32_t sort_tile, uint64_t sort_tiles, const ScanCtx& sc,
                            const Segs& seg, const RecordBufs& bufs, uint32_t lbits, bool packed, uint64_t* brick_key,
                            uint64_t* masks, uint32_t* first_dev, uint64_t nb) {
    constexpr uint32_t tile = kBlock * ITEMS;
    const uint32_t subs = (sort_tile + tile - 1) / tile;
    // brick-head look-back statuses follow the kept-record ones
    uint64_t* status_heads = sc.status + sort_tiles * subs;
    const dim3 grid((uint32_t)std::min<uint64_t>(sort_tiles * subs, 2048));
    if (packed)
        hipLaunchKernelGGL((k_unique_bricks<ITEMS, true>), grid, dim3(kBlock), 0, s, sc, status_heads, seg, subs, bufs,
                           lbits, brick_key, masks, first_dev, nb);
    else
        hipLaunchKernelGGL((k_unique_bricks<ITEMS, false>), grid, dim3(kBlock), 0, s, sc, status_heads, seg, subs,
                           bufs, lbits, brick_key, masks, first_dev, nb);
}

// unique + bricks with scan tiles of 256 * items records
void launch_unique_bricks(hipStream_t s, uint32_t items, uint32_t sort_tile, uint64_t sort_tiles, const ScanCtx& sc,
                          const Segs& seg, const RecordBufs& bufs, uint32_t lbits, bool packed, uint64_t* brick_key,
                          uint64_t* masks, uint32_t* first_dev, uint64_t nb) {
    switch (items) {
        case 12: launch_unique_bricks_t<12>(s, sort_tile, sort_tiles, sc, seg, bufs, lbits, packed, brick_key, masks, first_dev, nb); break;
        case 16: launch_unique_bricks_t<16>(s, sort_tile, sort_tiles, sc, seg, bufs, lbits, packed, brick_key, masks, first_dev, nb); break;
        case 20: launch_unique_bricks_t<20>(s, sort_tile, sort_tiles, sc, seg, bufs, lbits, packed, brick_key, masks, first_dev, nb); break;
        default: launch_unique_bricks_t<8>(s, sort_tile, sort_tiles, sc, seg, bufs, lbits, packed, brick_key, masks, first_dev, nb); break;
    }
}

}  // namespace

// A chunk-batch build captured into a HIP graph (pyrhip_graph_*): every
// buffer is allocated once for the largest batch, the kernels take their
// sizes from the device arrays (offsets, counters), and a replay ends with
// the counters, roots and first voxels copied into pinned host memory.
struct GraphBuild {
    std::vector<void*> scratch;
    uint64_t* kb[2] = {nullptr, nullptr};
    uint32_t* vb[2] = {nullptr, nullptr};
    uint64_t* leaves = nullptr;
    uint32_t* nodes = nullptr;
    uint32_t* host = nullptr;  // pinned: counters [C_TILE], roots [count], first [count]
    bool packed = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Whether a build of n records (chunk_count chunks, 0 = single) takes the
// bucket finish under the resolved tuning `tun`.
bool bucket_finish(const PyrhipBuildTuning& tun, uint32_t log2_size, uint64_t n, uint32_t chunk_count) {
    const uint32_t lbits = local_bits(log2_size);
    const uint32_t lb = std::min(kBucketBits, lbits);
    const uint64_t cap = (chunk_count ? (uint64_t)chunk_count : 1ull) << (lbits - lb);
    if (lbits - lb > 20 || cap > (1ull << 24)) return false;
    return tun.finish == 2 || (tun.finish == 0 && n >= (chunk_count ? 512ull : 2048ull) * cap);
}

// A build without allocation (pyrhip_build_static): every buffer, outputs
// included, is carved out of [base, base + size); the results are copied into
// the caller's buffers at the end.
struct FixedMemory {
    uint8_t* base = nullptr;
    uint64_t size = 0;
    const PyrhipBuildOutput* dst = nullptr;
};

// One build: single geometry (enc.chunk_count == 0) or chunk batch. With
// `gb` (batches only) the launches are captured into a graph on `s` instead
// of run, and the buffers move to gb. With `fixed` (single builds only) no
// memory is allocated.
PyrResult run_build(const EncodeArgs& enc_in, uint32_t log2_size, const PyrhipBuildTuning* tuning, hipStream_t s,
                    PyrhipBuilt* out, uint32_t* roots_host, uint32_t* first_host, GraphBuild* gb = nullptr,
                    const FixedMemory* fixed = nullptr) {
    memset(out, 0, sizeof(*out));
    EncodeArgs enc = enc_in;
    const uint64_t n64 = (uint64_t)enc.n_old + enc.n_new;
    if (n64 >= 0xFFFFFFFFull) return fail(PYR_ERROR_INVALID_ARGUMENT, "too many voxels for one build (%llu)",
                                          (unsigned long long)n64);
    const uint32_t n = (uint32_t)n64, nmax = std::max<uint32_t>(n, 1);
    // fields the caller leaves 0 come from the tuning installed for this
    // workload class (autotuner), else from the built-in defaults
    const PyrhipBuildTuning tun = resolve_tuning(tuning, log2_size, n, enc.chunk_count);
    const uint32_t rb = tun.radix_bits;
    if (rb < 8 || rb > 11) return fail(PYR_ERROR_INVALID_ARGUMENT, "radix_bits must be in [8, 11]");
    if (tun.pair_layout > 2) return fail(PYR_ERROR_INVALID_ARGUMENT, "pair_layout must be 0, 1 or 2");
    const uint32_t sort_items = tun.sort_items;
    if (sort_items != 8 && sort_items != 12 && sort_items != 16 && sort_items != 20 && sort_items != 24)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "sort_items must be 8, 12, 16, 20 or 24");
    const uint32_t scan_items = tun.scan_items;
    if (scan_items != 8 && scan_items != 12 && scan_items != 16 && scan_items != 20)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "scan_items must be 8, 12, 16 or 20");
    const uint32_t load_pct = tun.hash_load_pct;
    if (load_pct < 25 || load_pct > 95) return fail(PYR_ERROR_INVALID_ARGUMENT, "hash_load_pct must be in [25, 95]");
    if (tun.status_words < 1 || tun.status_words > 2)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "status_words must be 1 or 2");
    if (tun.counted_pass > 2) return fail(PYR_ERROR_INVALID_ARGUMENT, "counted_pass must be 1 or 2");
    if (tun.levels > 2) return fail(PYR_ERROR_INVALID_ARGUMENT, "levels must be 0, 1 or 2");
    const uint32_t lbits = local_bits(log2_size);
    // records sort by chunk-local key inside their chunk's segment: the key
    // is the local bits plus the invalid bit; the chunk index comes back into
    // the keys at the unique step (chunk << lbits | local, which must fit)
    if (lbits + enc.chunk_bits > 63) return fail(PYR_ERROR_INVALID_ARGUMENT, "too many chunks for this chunk size");
    const uint32_t key_bits = lbits + 1;
    // bucket finish: sort the bits from lb up only, then one key table per
    // bucket of 2^lb keys; by default when a single build's buckets hold 2048
    // records or more on average (a batch: 512 per bucket of its largest chunks)
    if (tun.finish > 2) return fail(PYR_ERROR_INVALID_ARGUMENT, "finish must be 0, 1 or 2");
    const uint32_t lb_try = std::min(kBucketBits, lbits);
    const uint64_t bucket_cap = (enc.chunk_count ? (uint64_t)enc.chunk_count : 1ull) << (lbits - lb_try);
    const bool bucket_ok = lbits - lb_try <= 20 && bucket_cap <= (1ull << 24);
    if (tun.finish == 2 && !bucket_ok)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "the bucket finish needs at most 2^24 buckets");
    const bool bucketed = bucket_finish(tun, log2_size, n, enc.chunk_count);
    const uint32_t lb = bucketed ? lb_try : 0;
    const uint32_t buckets = bucketed ? (uint32_t)bucket_cap : 0;  // a batch's is an upper bound
    const uint32_t passes = (key_bits - lb + rb - 1) / rb;
    // keys of at most 32 bits travel packed with their attribute in one u64
    // (16 instead of 24 bytes per record and pass)
    const bool packed = key_bits <= 32 && rb == 8 && tun.pair_layout != 1;
    const uint32_t radix = 1u << rb;
    const uint64_t k = enc.chunk_count ? enc.chunk_count : 1;

    // upper bounds (gpu_build.zig:142-154, 291-300)
    auto cells = [&](uint32_t lv) -> uint64_t {  // cells of edge 2^lv in all chunks
        uint64_t per = 1ull << (3 * (log2_size - std::min(lv, log2_size)));
        return std::max<uint64_t>(std::min<uint64_t>(nmax, per * k), 1);
    };
    const uint64_t nb = cells(2);  // bricks, at most: the unique step writes this many brick slots
    // Sizes of the leaf and level buffers for b bricks: parents of bricks,
    // hash capacity, node words (each level's parents, at most one per child
    // and one per cell of the level above, take 2 words plus one per child).
    // The brick count is an upper bound from the record count unless the
    // build reads the real count back after the unique step.
    struct LevelSizes {
        uint64_t nb, np2, cap, words;
    };
    auto level_sizes = [&](uint64_t b) {
        LevelSizes z;
        z.nb = std::max<uint64_t>(b, 1);
        z.np2 = std::min(z.nb, cells(3));
        z.cap = hash_cap_for(z.nb, load_pct);
        z.words = 2;
        uint64_t c = z.nb;
        for (uint32_t lv = 2; lv < log2_size; lv++) {
            const uint64_t par = std::min(c, cells(lv + 1));
            z.words += 2 * par + c;
            c = par;
        }
        return z;
    };
    // bytes of the buffers sized by the brick count (table + owner, slot,
    // brick elements, parent elements and tuples, leaves, node words)
    auto level_bytes = [](const LevelSizes& z) {
        return z.cap * 12 + z.nb * (4 + 12 + 8) + z.np2 * (20 + (uint64_t)kNodeStride * 4 + 4) + z.words * 4;
    };
    LevelSizes lz = level_sizes(nb);
    // when the bounds would hold more than 8 bytes per record plus 1 MiB for
    // the brick-sized buffers (sparse inputs with about one brick per record,
    // batches whose chunk cells far exceed their bricks) the build reads the
    // brick count back once after the unique step and sizes them exactly, so
    // a build never holds more than that beyond its real need; smaller
    // builds skip the read-back; graph captures and the static scratch keep
    // the bounds
    const bool exact_levels = !gb && !fixed && level_bytes(lz) > 8ull * nmax + (1ull << 20);

    const auto host_t0 = std::chrono::steady_clock::now();
    Alloc A{{}, s};
    if (fixed) {
        A.base = fixed->base;
        A.size = fixed->size;
    }
    uint32_t* counters = A.get<uint32_t>(C_COUNT);
    const uint32_t segs = enc.chunk_count ? enc.chunk_count : 1;
    const uint64_t hist_words = (uint64_t)passes * segs * radix;
    uint32_t* hist = A.get<uint32_t>(hist_words * 2);
    uint32_t* digit_base = hist + hist_words;
    const uint32_t sort_tile = kBlock * sort_items;
    // record tiles: a single build has exactly ceil(n / tile); a batch's are
    // counted on the device, at most one partial tile more per chunk
    const uint64_t sort_tiles = (nmax + sort_tile - 1) / sort_tile + (enc.chunk_count ? enc.chunk_count : 0);
    if (sort_tiles * sort_tile >= 0xFFFFFFFFull)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "too many voxels and chunks for one build");
    Segs seg{};
    seg.single = enc.chunk_count == 0;
    seg.count = segs;
    seg.n = n;
    seg.tile = sort_tile;
    seg.off = enc.offsets;
    uint32_t* tile_first = enc.chunk_count ? A.get<uint32_t>(enc.chunk_count + 1) : nullptr;
    seg.desc = enc.chunk_count ? A.get<uint4>(sort_tiles) : nullptr;
    seg.tiles = counters + C_TILES;
    const uint64_t scan_tiles = (nmax + kScanTile - 1) / kScanTile;
    // sort passes: one word per (tile, digit); unique: one per kScanTile piece of a tile
    // (sort passes: one word per tile and digit; unique/bricks: two per scan piece)
    // packed records in segments of fewer than 2^30 records: narrow sort
    // look-back words in an array of their own, cleared before every pass
    const bool narrow = packed && n < (1u << 30) && tun.status_words == 2;
    // (bucket finish: two words per bucket)
    const uint64_t status_words = std::max<uint64_t>(
        std::max<uint64_t>(sort_tiles * std::max<uint32_t>(narrow ? 0u : radix,
                                                            2 * ((sort_tile + kScanTile - 1) / kScanTile)),
                           scan_tiles),
        2ull * buckets);
    uint64_t* status = A.get<uint64_t>(status_words);
    uint32_t* bucket_start = bucketed ? A.get<uint32_t>(buckets) : nullptr;
    uint32_t* bucket_end = bucketed ? A.get<uint32_t>(buckets) : nullptr;
    uint32_t* bucket_seg = bucketed && enc.chunk_count ? A.get<uint32_t>(buckets) : nullptr;
    uint32_t* first_bucket = bucketed && enc.chunk_count ? A.get<uint32_t>(enc.chunk_count + 1) : nullptr;
    uint32_t* status32 = narrow ? A.get<uint32_t>(sort_tiles * radix) : nullptr;
    // pass 0 counted by encode (8-bit digits): per-tile run prefixes and per-run totals
    const uint32_t eblocks = (uint32_t)std::min<uint64_t>(sort_tiles, 2048);
    const uint32_t per_block = (uint32_t)((sort_tiles + eblocks - 1) / eblocks);
    const bool counted = radix == (uint32_t)kBlock && tun.counted_pass != 2;
    uint32_t* tile_run = counted ? A.get<uint32_t>(sort_tiles * radix) : nullptr;
    uint32_t* run_sum = counted ? A.get<uint32_t>((uint64_t)eblocks * radix * 2) : nullptr;
    uint64_t* brick_key = A.get<uint64_t>(nb);
    uint64_t* masks = A.get<uint64_t>(nb);
    uint64_t* table = nullptr;
    uint32_t* owner = nullptr;
    uint32_t* slot = nullptr;
    LevelArrays set[2];
    uint32_t* node_tmp = nullptr;
    uint32_t* node_len = nullptr;
    uint32_t* roots_dev = enc.chunk_count ? A.get<uint32_t>(enc.chunk_count * 2) : nullptr;
    uint32_t* first_dev = enc.chunk_count ? roots_dev + enc.chunk_count : nullptr;
    // caller-owned outputs; of the two key/value buffer pairs one ends up
    // holding the unique voxels (Built.keys/attrs), the other is freed
    uint64_t* kb[2] = {nullptr, nullptr};
    uint32_t* vb[2] = {nullptr, nullptr};
    uint64_t* leaves_out = nullptr;
    uint32_t* nodes_out = nullptr;
    bool out_ok = true;
    auto out_alloc = [&](void** p, uint64_t bytes) -> bool {
        if (fixed) return (*p = A.raw(bytes)) != nullptr;
        A.requested += std::max<uint64_t>(bytes, 256);
        return pool_alloc(p, bytes, s) == hipSuccess;
    };
    for (int q = 0; q < 2; q++) {
        out_ok &= out_alloc((void**)&kb[q], (uint64_t)nmax * 8);
        // packed records need one value array (for the unique attributes)
        if (q == 0 || !packed) out_ok &= out_alloc((void**)&vb[q], (uint64_t)nmax * 4);
    }
    // the buffers sized by the brick count, the leaves and node words among them
    auto alloc_levels = [&]() {
        bool ok = true;
        table = A.get<uint64_t>(lz.cap);
        owner = A.get<uint32_t>(lz.cap);
        slot = A.get<uint32_t>(lz.nb);
        for (int q = 0; q < 2; q++) {
            set[q].key = q == 0 ? brick_key : A.get<uint64_t>(lz.np2);
            set[q].rep = A.get<uint32_t>(q == 0 ? lz.nb : lz.np2);
            set[q].pos = A.get<uint32_t>(q == 0 ? lz.nb : lz.np2);
            set[q].count = A.get<uint32_t>(q == 0 ? lz.nb : lz.np2);
        }
        node_tmp = A.get<uint32_t>(lz.np2 * kNodeStride);
        node_len = A.get<uint32_t>(lz.np2);
        ok &= out_alloc((void**)&leaves_out, lz.nb * 8);
        ok &= out_alloc((void**)&nodes_out, lz.words * 4);
        return ok && !A.failed;
    };
    if (!exact_levels) out_ok &= alloc_levels();
    auto free_outputs = [&]() {
        if (fixed) return;
        for (int q = 0; q < 2; q++) {
            pool_free(kb[q], s);
            pool_free(vb[q], s);
        }
        pool_free(leaves_out, s);
        pool_free(nodes_out, s);
    };
    if (fixed && (A.failed || !out_ok))
        return fail(PYR_ERROR_CAPACITY, "a build of %u voxels at log2_size %u needs %llu bytes of scratch; the static scratch holds %llu",
                    n, log2_size, (unsigned long long)A.used, (unsigned long long)A.size);
    if (A.failed || !out_ok) {
        A.release();
        free_outputs();
        (void)hipGetLastError();
        return fail(PYR_ERROR_OUT_OF_MEMORY, "device memory for a build of %u voxels", n);
    }

    const auto host_t1 = std::chrono::steady_clock::now();  // allocations done
    // (a graph replays on other streams: the waits that reused cached blocks
    // put on `s` must be over before capture)
    if (gb && (hipStreamSynchronize(s) != hipSuccess || hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)) {
        (void)hipGetLastError();
        A.release();
        free_outputs();
        return fail(PYR_ERROR_CUDA, "graph capture could not start");
    }
    Timer tm;
    tm.s = s;
    if (gb || fixed) tm.on = false;
    tm.mark();
    // zero state
    fill32(s, counters, 0, C_COUNT);
    fill32(s, hist, 0, hist_words);
    fill64(s, status, 0, status_words);
    if (!bucketed) fill64(s, masks, 0, nb);  // (the bucket finish writes whole masks)
    if (enc.chunk_count) fill32(s, roots_dev, 0xFFFFFFFFu, enc.chunk_count * 2);

    uint32_t launch = 0, epoch = 0;
    auto tile_ctr = [&]() { return counters + C_TILE + (launch++ % kMaxLaunches); };

    // 1. encode + histograms
    enc.keys = kb[0];
    enc.vals = vb[0];
    enc.hist = hist;
    enc.passes = passes;
    enc.radix_bits = rb;
    enc.shift0 = lb;
    enc.packed = packed;
    if (n) {
        if (enc.chunk_count) {
            hipLaunchKernelGGL(k_seg_tiles, dim3(1), dim3(kBlock), 0, s, enc.offsets, enc.chunk_count, n,
                               gb ? 1u : 0u, sort_tile, tile_first, counters + C_TILES, counters + C_ERR);
            hipLaunchKernelGGL(k_seg_desc, dim3(grid_for(enc.chunk_count, kBlock / kWave, 1024)), dim3(kBlock), 0, s,
                               enc.offsets, tile_first, enc.chunk_count, sort_tile, counters + C_TILES,
                               (uint4*)seg.desc);
        }
        enc.tile_run = tile_run;
        enc.run_sum = run_sum;
        hipLaunchKernelGGL(k_encode, dim3(eblocks), dim3(kBlock), passes * radix * sizeof(uint32_t), s, enc, seg,
                           per_block);
        if (counted)
            hipLaunchKernelGGL(k_run_scan, dim3(kBlock), dim3(kBlock), 0, s, run_sum, run_sum + (uint64_t)eblocks * radix,
                               eblocks);
        hipLaunchKernelGGL(k_digit_scan, dim3(passes * segs), dim3(kBlock), 0, s, hist, digit_base, seg, passes,
                           radix, counters);
        hipLaunchKernelGGL(k_pass_plan, dim3(1), dim3(kWave), 0, s, passes, counters);
    }
    const RecordBufs bufs{{kb[0], kb[1]}, {vb[0], vb[1]}};
    tm.mark();
    // 2. radix passes
    // (the device skips passes that move nothing and picks each pass's
    // source buffer, so the buffers are named by index here)
    for (uint32_t p = 0; p < passes && n; p++) {
        // pass 0 reads the records in input order, counted by encode; later
        // passes look back (counting them too was measured slower: their
        // tiles hold most digits, and tiles that write whenever they are ready
        // scatter their short runs over the whole output at once, while the
        // look-back keeps the write-outs of neighbouring tiles together)
        const bool use_counts = counted && p == 0;
        if (narrow && !use_counts) fill32(s, status32, 0, sort_tiles * radix);
        SortArgs sa{bufs, p, counters, seg, (packed ? 32 : 0) + lb + p * rb, digit_base + (uint64_t)p * segs * radix,
                    status, status32, ++epoch, tile_ctr(), counters + C_ERR, bucketed ? kSortPeel : 0u,
                    use_counts ? tile_run : nullptr, use_counts ? run_sum + (uint64_t)eblocks * radix : nullptr,
                    per_block};
        launch_sort(s, rb, sort_items, packed, (uint32_t)sort_tiles, sa);
    }
    tm.mark();
    // 3. unique into the buffer pair that does not hold the sorted records;
    // it becomes Built.keys/attrs
    if (n && bucketed) {
        ScanCtx sc{status, ++epoch, tile_ctr(), counters};
        if (enc.chunk_count)
            hipLaunchKernelGGL(k_bucket_plan, dim3(1), dim3(kBlock), 0, s, enc.chunk_log2, enc.chunk_count, log2_size,
                               lb, first_bucket, counters);
        const dim3 bgrid(grid_for(buckets, kBlock, 1024));
        const dim3 ugrid(std::min<uint32_t>(buckets, 1024));
        if (packed) {
            hipLaunchKernelGGL(k_bucket_bounds<true>, bgrid, dim3(kBlock), 0, s, counters, bufs, seg, first_bucket, lb,
                               buckets, bucket_start, bucket_end, bucket_seg);
            hipLaunchKernelGGL(k_bucket_unique<true>, ugrid, dim3(kBucketBlock), 0, s, sc, status + buckets,
                               bucket_start, bucket_end, bucket_seg, buckets, bufs, lb, lbits, brick_key, masks,
                               first_dev, nb, n);
        } else {
            hipLaunchKernelGGL(k_bucket_bounds<false>, bgrid, dim3(kBlock), 0, s, counters, bufs, seg, first_bucket,
                               lb, buckets, bucket_start, bucket_end, bucket_seg);
            hipLaunchKernelGGL(k_bucket_unique<false>, ugrid, dim3(kBucketBlock), 0, s, sc, status + buckets,
                               bucket_start, bucket_end, bucket_seg, buckets, bufs, lb, lbits, brick_key, masks,
                               first_dev, nb, n);
        }
    } else if (n) {
        ScanCtx sc{status, ++epoch, tile_ctr(), counters};
        launch_unique_bricks(s, scan_items, sort_tile, sort_tiles, sc, seg, bufs, lbits, packed, brick_key, masks,
                             first_dev, nb);
    }
    tm.mark();
    // the real brick count, when the brick-sized buffers are sized by it
    if (exact_levels) {
        uint32_t bricks = 0;
        hipError_t be = hipMemcpyAsync(&bricks, counters + C_B, sizeof(bricks), hipMemcpyDeviceToHost, s);
        if (be == hipSuccess) be = hipStreamSynchronize(s);
        if (be != hipSuccess) {
            (void)hipGetLastError();
            A.release();
            free_outputs();
            return fail(PYR_ERROR_CUDA, "build: %s", hipGetErrorString(be));
        }
        lz = level_sizes(std::min<uint64_t>(bricks, nb));  // (a failed unique step may report anything)
        if (!alloc_levels()) {
            A.release();
            free_outputs();
            (void)hipGetLastError();
            return fail(PYR_ERROR_OUT_OF_MEMORY, "device memory for a build of %u voxels", n);
        }
    }
    // 4. leaves
    const uint32_t brick_grid = grid_for(lz.nb, kScanTile, 2048);
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_for(lz.cap, 1024, 2048)), dim3(kBlock), 0, s, counters, (uint32_t)C_B,
                       (void*)table, 8u, owner, lz.cap, load_pct);
    hipLaunchKernelGGL(k_leaf_insert, dim3(grid_for(lz.nb, kBlock * 4, kInsertBlocks)), dim3(kBlock), 0, s, counters, masks,
                       table, owner, slot, set[0].count, lz.cap, (uint32_t)lz.nb, load_pct);
    {
        ScanCtx sc{status, ++epoch, tile_ctr(), counters};
        hipLaunchKernelGGL(k_leaf_compact, dim3(brick_grid), dim3(kBlock), 0, s, sc, masks, owner, slot,
                           leaves_out, set[0].rep, set[0].pos, (uint32_t)lz.nb);
    }
    tm.mark();
    // 5. node levels: one workgroup and one launch for small builds, else
    // four launches per level
    int ch = 0;
    const uint64_t set_cap[2] = {lz.nb, lz.np2};
    const bool small_levels = tun.levels == 2 || (tun.levels == 0 && lz.nb <= kLevelsSmallBricks);
    if (small_levels && log2_size > 2) {
        LevelArgs la{};
        la.chunk_log2 = enc.chunk_log2;
        la.roots = roots_dev;
        la.ch = set[0];
        la.par = set[1];
        la.node_tmp = node_tmp;
        la.node_len = node_len;
        la.table = (uint32_t*)table;
        la.owner = owner;
        la.slot = slot;
        la.table_cap = lz.cap;
        la.word_cap = lz.words;
        la.chunk_count = enc.chunk_count;
        la.load_pct = load_pct;
        hipLaunchKernelGGL(k_levels_small, dim3(1), dim3(kLevelsBlock), 0, s, counters, la, log2_size,
                           (uint32_t)set_cap[0], (uint32_t)set_cap[1], (uint32_t)lz.np2, nodes_out);
        ch = (log2_size - 2) & 1;
    }
    for (uint32_t lv = 2; lv < log2_size && !small_levels; lv++) {
        LevelArgs la{};
        la.lv = lv;
        la.top_shift = 3 * (log2_size - lv);
        la.chunk_log2 = enc.chunk_log2;
        la.roots = roots_dev;
        la.ch = set[ch];
        la.par = set[ch ^ 1];
        la.node_tmp = node_tmp;
        la.node_len = node_len;
        la.table = (uint32_t*)table;
        la.owner = owner;
        la.slot = slot;
        la.ch_cap = (uint32_t)set_cap[ch];
        la.par_cap = (uint32_t)std::min(set_cap[ch ^ 1], lz.np2);
        la.table_cap = lz.cap;
        la.word_cap = lz.words;
        la.chunk_count = enc.chunk_count;
        la.load_pct = load_pct;
        const uint64_t nch = lv == 2 ? lz.nb : std::min(lz.nb, cells(lv));
        const uint32_t g = grid_for(nch, kScanTile, 2048);
        hipLaunchKernelGGL(k_hash_clear, dim3(grid_for(hash_cap_for(nch, load_pct), 1024, 2048)), dim3(kBlock), 0, s,
                           counters, (uint32_t)(C_NCH + lv), (void*)table, 4u, owner, lz.cap, load_pct);
        ScanCtx sc{status, ++epoch, tile_ctr(), counters};
        hipLaunchKernelGGL(k_level_build, dim3(g), dim3(kBlock), 0, s, sc, la);
        hipLaunchKernelGGL(k_level_insert, dim3(grid_for(std::min(lz.nb, cells(lv + 1)), kBlock * 4, kInsertBlocks)),
                           dim3(kBlock), 0, s, counters, la);
        ScanCtx sc2{status, ++epoch, tile_ctr(), counters};
        hipLaunchKernelGGL(k_level_compact, dim3(grid_for(std::min(lz.nb, cells(lv + 1)), kScanTile, 2048)),
                           dim3(kBlock), 0, s, sc2, la, nodes_out);
        ch ^= 1;
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s, counters, log2_size, set[ch], roots_dev,
                       nodes_out, (uint32_t)std::min(set_cap[ch], lz.np2), enc.chunk_count, lz.words);
    tm.mark();

    // single read-back (gpu_build.zig:338-344); a graph copies into its pinned buffer
    uint32_t c[C_TILE];
    uint32_t* const hc = gb ? gb->host : c;
    if (gb) {
        roots_host = gb->host + C_TILE;
        first_host = roots_host + enc.chunk_count;
    }
    hipError_t e = hipMemcpyAsync(hc, counters, sizeof(c), hipMemcpyDeviceToHost, s);
    if (enc.chunk_count && e == hipSuccess) {
        e = hipMemcpyAsync(roots_host, roots_dev, enc.chunk_count * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(first_host, first_dev, enc.chunk_count * 4, hipMemcpyDeviceToHost, s);
    }
    if (gb) {
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = ce;
        hipGraphExec_t exec = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            A.release();
            free_outputs();
            return fail(PYR_ERROR_CUDA, "graph capture: %s", hipGetErrorString(e));
        }
        gb->graph = graph;
        gb->exec = exec;
        gb->scratch.swap(A.ptrs);
        for (int q = 0; q < 2; q++) {
            gb->kb[q] = kb[q];
            gb->vb[q] = vb[q];
        }
        gb->leaves = leaves_out;
        gb->nodes = nodes_out;
        gb->packed = packed;
        return PYR_OK;
    }
    A.release();
    const auto host_t2 = std::chrono::steady_clock::now();  // everything enqueued
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const auto host_t3 = std::chrono::steady_clock::now();
    if (e == hipSuccess) e = hipGetLastError();
    if (g_timing && !fixed && e == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_times_mutex);
        g_times.encode_ms = tm.between(0, 1);
        g_times.sort_ms = tm.between(1, 2);
        g_times.unique_ms = tm.between(2, 3);
        g_times.leaves_ms = tm.between(3, 4);
        g_times.levels_ms = tm.between(4, 5);
        g_times.total_ms = tm.between(0, 5);
        g_times.sort_passes = (uint32_t)__builtin_popcount(c[C_PASS]);  // passes that moved records
        g_times.radix_bits = rb;
        g_times.record_bytes = packed ? 8 : 12;
        g_times.sort_items = sort_items;
        g_times.status_bytes = narrow ? 4 : 8;
        g_times.finish = bucketed ? 2 : 1;
        g_times.device_bytes = A.requested;
        using ms = std::chrono::duration<float, std::milli>;
        g_times.host_alloc_ms = ms(host_t1 - host_t0).count();
        g_times.host_enqueue_ms = ms(host_t2 - host_t1).count();
        g_times.host_wait_ms = ms(host_t3 - host_t2).count();
    }
    if (e != hipSuccess || c[C_ERR] != 0) {
        free_outputs();
        if (e != hipSuccess) return fail(PYR_ERROR_CUDA, "build: %s", hipGetErrorString(e));
        return fail(PYR_ERROR_CUDA, "build: device error flags 0x%x", c[C_ERR]);
    }
    if (fixed) {
        // results into the caller's buffers (the scratch keeps them until the
        // next call on this device)
        const PyrhipBuildOutput& d = *fixed->dst;
        const uint32_t fin = c[C_FINAL] & 1;
        const uint64_t* keys = kb[fin ^ 1];
        const uint32_t* attrs = packed ? vb[0] : vb[fin ^ 1];
        if (c[C_WORDS] > d.node_capacity || c[C_LEAVES] > d.leaf_capacity || c[C_U] > d.voxel_capacity)
            return fail(PYR_ERROR_CAPACITY, "build output of %u node words, %u leaves, %u voxels exceeds the buffers (%llu, %llu, %llu)",
                        c[C_WORDS], c[C_LEAVES], c[C_U], (unsigned long long)d.node_capacity,
                        (unsigned long long)d.leaf_capacity, (unsigned long long)d.voxel_capacity);
        e = hipMemcpyAsync((void*)(uintptr_t)d.nodes, nodes_out, (uint64_t)c[C_WORDS] * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && c[C_LEAVES])
            e = hipMemcpyAsync((void*)(uintptr_t)d.leaves, leaves_out, (uint64_t)c[C_LEAVES] * 8,
                               hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && c[C_U] && d.keys)
            e = hipMemcpyAsync((void*)(uintptr_t)d.keys, keys, (uint64_t)c[C_U] * 8, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && c[C_U] && d.attrs)
            e = hipMemcpyAsync((void*)(uintptr_t)d.attrs, attrs, (uint64_t)c[C_U] * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(PYR_ERROR_CUDA, "build output copy: %s", hipGetErrorString(e));
        }
        out->log2_size = log2_size;
        out->root = c[C_ROOT];
        out->nodes = d.nodes;
        out->node_words = c[C_WORDS];
        out->leaves = d.leaves;
        out->leaf_count = c[C_LEAVES];
        out->keys = d.keys;
        out->attrs = d.attrs;
        out->voxel_count = c[C_U];
        return PYR_OK;
    }
    out->log2_size = log2_size;
    out->root = c[C_ROOT];
    out->nodes = (uint64_t)(uintptr_t)nodes_out;
    out->node_words = c[C_WORDS];
    out->leaves = (uint64_t)(uintptr_t)leaves_out;
    out->leaf_count = c[C_LEAVES];
    const uint32_t fin = c[C_FINAL] & 1;  // buffer pair of the sorted records
    out->keys = (uint64_t)(uintptr_t)kb[fin ^ 1];
    out->attrs = (uint64_t)(uintptr_t)(packed ? vb[0] : vb[fin ^ 1]);
    out->voxel_count = c[C_U];
    pool_free(kb[fin], s);
    if (!packed) pool_free(vb[fin], s);
    return PYR_OK;
}

}  // namespace pyr

// Scratch of pyrhip_build_static: every buffer of such a build (records,
// histograms, look-back words, hash tables, level arrays, outputs before the
// copy) is carved out of this array, one copy per device, held until the
// process ends. In host code `scratch` names a host-side stand-in, not device
// memory: the device address comes from hipGetSymbolAddress
// (static_scratch_address) and only that address is passed to kernels.
constexpr uint64_t kStaticScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

std::mutex g_static_mutex;
std::vector<uint8_t*> g_static_ptr;  // per device; null until looked up

// Device address of the current device's copy of `scratch`, looked up once
// per device (with that device current).
PyrResult static_scratch_address(uint8_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return pyr::fail(PYR_ERROR_CUDA, "hipGetDevice: %s", hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_static_mutex);
    if ((size_t)dev >= g_static_ptr.size()) g_static_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_static_ptr[(size_t)dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return pyr::fail(PYR_ERROR_CUDA, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        }
        g_static_ptr[(size_t)dev] = (uint8_t*)p;
    }
    *out = g_static_ptr[(size_t)dev];
    return PYR_OK;
}

}  // namespace

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_build(uint32_t log2_size, uint64_t old_keys, uint64_t old_vals, uint32_t n_old,
                               uint64_t voxels, uint32_t n_new, const PyrhipBuildTuning* tuning, void* stream,
                               PyrhipBuilt* out) {
    diag_clear();
    if (!out) return fail(PYR_ERROR_INVALID_ARGUMENT, "out is null");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if ((n_old && (!old_keys || !old_vals)) || (n_new && !voxels))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null input with a non-zero count");
    EncodeArgs enc{};
    enc.old_keys = (const uint64_t*)(uintptr_t)old_keys;
    enc.old_vals = (const uint32_t*)(uintptr_t)old_vals;
    enc.n_old = n_old;
    enc.voxels = (const int4*)(uintptr_t)voxels;
    enc.n_new = n_new;
    enc.log2_size = log2_size;
    return run_build(enc, log2_size, tuning, (hipStream_t)stream, out, nullptr, nullptr);
}

PYR_API PyrResult pyrhip_build_static(uint32_t log2_size, uint64_t old_keys, uint64_t old_vals, uint32_t n_old,
                                      uint64_t voxels, uint32_t n_new, const PyrhipBuildTuning* tuning, void* stream,
                                      const PyrhipBuildOutput* dst, PyrhipBuilt* out) {
    diag_clear();
    if (!out || !dst) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (!dst->nodes || !dst->leaves) return fail(PYR_ERROR_INVALID_ARGUMENT, "null output buffer");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if ((n_old && (!old_keys || !old_vals)) || (n_new && !voxels))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null input with a non-zero count");
    uint8_t* base = nullptr;
    const PyrResult r = static_scratch_address(&base);
    if (r != PYR_OK) return r;
    EncodeArgs enc{};
    enc.old_keys = (const uint64_t*)(uintptr_t)old_keys;
    enc.old_vals = (const uint32_t*)(uintptr_t)old_vals;
    enc.n_old = n_old;
    enc.voxels = (const int4*)(uintptr_t)voxels;
    enc.n_new = n_new;
    enc.log2_size = log2_size;
    FixedMemory fm;
    fm.base = base;
    fm.size = kStaticScratchBytes;
    fm.dst = dst;
    return run_build(enc, log2_size, tuning, (hipStream_t)stream, out, nullptr, nullptr, nullptr, &fm);
}

PYR_API uint64_t pyrhip_build_static_limit(void) { return kStaticScratchBytes; }

PYR_API PyrResult pyrhip_build_chunks(uint32_t log2_size, const PyrhipChunkBatch* batch,
                                      const PyrhipBuildTuning* tuning, void* stream, PyrhipBuilt* out,
                                      uint32_t* roots, uint32_t* first_voxel) {
    diag_clear();
    if (!out || !batch || !roots || !first_voxel) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if (batch->count == 0) return fail(PYR_ERROR_INVALID_ARGUMENT, "empty chunk batch");
    if (!batch->offsets || (batch->total && !batch->voxels))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null batch buffers");
    EncodeArgs enc{};
    enc.voxels = (const int4*)(uintptr_t)batch->voxels;
    enc.n_new = batch->total;
    enc.log2_size = log2_size;
    enc.chunk_count = batch->count;
    enc.capacity = batch->capacity;
    enc.offsets = (const uint32_t*)(uintptr_t)batch->offsets;
    enc.chunk_log2 = (const uint8_t*)(uintptr_t)batch->chunk_log2;
    uint32_t bits = 0;
    while ((1u << bits) < batch->count) bits++;
    enc.chunk_bits = bits;
    return run_build(enc, log2_size, tuning, (hipStream_t)stream, out, roots, first_voxel);
}

struct PyrhipGraph {
    pyr::GraphBuild gb;
    uint32_t log2_size = 0, count = 0, capacity = 0, total_max = 0;
    bool mixed = false;
    hipStream_t stream = nullptr;  // the capture stream
    void* voxels = nullptr;        // inputs the graph reads
    uint32_t* offsets = nullptr;
    uint8_t* chunk_log2 = nullptr;
    // pinned staging of the caller's offsets and chunk sizes for each replay
    uint32_t* host_offsets = nullptr;
    uint8_t* host_log2 = nullptr;
};

PYR_API void pyrhip_graph_destroy(PyrhipGraph* g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->gb.exec) (void)hipGraphExecDestroy(g->gb.exec);
    if (g->gb.graph) (void)hipGraphDestroy(g->gb.graph);
    for (void* p : g->gb.scratch) pool_free(p, g->stream);
    for (int q = 0; q < 2; q++) {
        pool_free(g->gb.kb[q], g->stream);
        pool_free(g->gb.vb[q], g->stream);
    }
    pool_free(g->gb.leaves, g->stream);
    pool_free(g->gb.nodes, g->stream);
    if (g->gb.host) (void)hipHostFree(g->gb.host);
    if (g->host_offsets) (void)hipHostFree(g->host_offsets);
    if (g->host_log2) (void)hipHostFree(g->host_log2);
    (void)hipFree(g->voxels);
    (void)hipFree(g->offsets);
    (void)hipFree(g->chunk_log2);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    (void)hipGetLastError();
    delete g;
}

PYR_API PyrResult pyrhip_graph_create(uint32_t log2_size, uint32_t count, uint32_t capacity, uint32_t total_max,
                                      int32_t mixed_sizes, const PyrhipBuildTuning* tuning, PyrhipGraph** out) {
    diag_clear();
    if (!out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if (count == 0 || total_max == 0) return fail(PYR_ERROR_INVALID_ARGUMENT, "empty graph batch");
    if (capacity && (uint64_t)count * capacity < total_max)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "total_max exceeds count * capacity");
    const uint64_t records = capacity ? (uint64_t)count * capacity : total_max;
    PyrhipGraph* g = new (std::nothrow) PyrhipGraph();
    if (!g) return fail(PYR_ERROR_OUT_OF_MEMORY, "graph");
    g->log2_size = log2_size;
    g->count = count;
    g->capacity = capacity;
    g->total_max = total_max;
    g->mixed = mixed_sizes != 0;
    bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&g->voxels, records * sizeof(PyrVoxel)) == hipSuccess;
    ok = ok && hipMalloc((void**)&g->offsets, ((uint64_t)count + 1) * 4) == hipSuccess;
    ok = ok && (!g->mixed || hipMalloc((void**)&g->chunk_log2, count) == hipSuccess);
    ok = ok && hipHostMalloc((void**)&g->gb.host, ((uint64_t)C_TILE + 2ull * count) * 4, 0) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&g->host_offsets, ((uint64_t)count + 1) * 4, 0) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&g->host_log2, count, 0) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        pyrhip_graph_destroy(g);
        return fail(PYR_ERROR_OUT_OF_MEMORY, "graph buffers for %u chunks and %u voxels", count, total_max);
    }
    EncodeArgs enc{};
    enc.voxels = (const int4*)g->voxels;
    enc.n_new = total_max;  // sizes for the buffers; the kernels read the offsets
    enc.log2_size = log2_size;
    enc.chunk_count = count;
    enc.capacity = capacity;
    enc.offsets = g->offsets;
    enc.chunk_log2 = g->chunk_log2;
    uint32_t bits = 0;
    while ((1u << bits) < count) bits++;
    enc.chunk_bits = bits;
    PyrhipBuilt unused{};
    const PyrResult r = run_build(enc, log2_size, tuning, g->stream, &unused, nullptr, nullptr, &g->gb);
    if (r != PYR_OK) {
        pyrhip_graph_destroy(g);
        return r;
    }
    *out = g;
    return PYR_OK;
}

PYR_API uint64_t pyrhip_graph_voxels(PyrhipGraph* g) { return g ? (uint64_t)(uintptr_t)g->voxels : 0; }

PYR_API PyrResult pyrhip_graph_launch(PyrhipGraph* g, uint64_t voxels, const uint32_t* offsets,
                                      const uint8_t* chunk_log2, void* stream, PyrhipBuilt* out, uint32_t* roots,
                                      uint32_t* first_voxel) {
    diag_clear();
    if (!g || !offsets || !out || !roots || !first_voxel) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (g->mixed != (chunk_log2 != nullptr))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "chunk_log2 must be given exactly for graphs of mixed sizes");
    // the kernels trust the offsets: check them here
    if (offsets[0] != 0) return fail(PYR_ERROR_INVALID_ARGUMENT, "offsets[0] must be 0");
    for (uint32_t c = 0; c < g->count; c++) {
        if (offsets[c + 1] < offsets[c]) return fail(PYR_ERROR_INVALID_ARGUMENT, "offsets decrease at chunk %u", c);
        if (g->capacity && offsets[c + 1] - offsets[c] > g->capacity)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "chunk %u exceeds the capacity", c);
        if (chunk_log2 && (chunk_log2[c] < 3 || chunk_log2[c] > g->log2_size))
            return fail(PYR_ERROR_INVALID_ARGUMENT, "chunk_log2[%u] outside [3, log2_size]", c);
    }
    const uint32_t total = offsets[g->count];
    if (total > g->total_max) return fail(PYR_ERROR_INVALID_ARGUMENT, "%u voxels, the graph holds %u", total, g->total_max);
    // The previous replay has finished (launches synchronise), so the pinned
    // staging buffers are free to take this replay's offsets and sizes.
    // (Replaying on the capture stream, ordered after the caller's by events,
    // added 0.17 ms to the server's median latency and was dropped.)
    const hipStream_t s = (hipStream_t)stream;
    memcpy(g->host_offsets, offsets, ((uint64_t)g->count + 1) * 4);
    if (chunk_log2) memcpy(g->host_log2, chunk_log2, g->count);
    hipError_t e = hipMemcpyAsync(g->offsets, g->host_offsets, ((uint64_t)g->count + 1) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && chunk_log2)
        e = hipMemcpyAsync(g->chunk_log2, g->host_log2, g->count, hipMemcpyHostToDevice, s);
    const uint64_t bytes = (g->capacity ? (uint64_t)g->count * g->capacity : total) * sizeof(PyrVoxel);
    if (e == hipSuccess && voxels && voxels != (uint64_t)(uintptr_t)g->voxels && bytes)
        e = hipMemcpyAsync(g->voxels, (const void*)(uintptr_t)voxels, bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipGraphLaunch(g->gb.exec, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_CUDA, "graph launch: %s", hipGetErrorString(e));
    }
    const uint32_t* c = g->gb.host;
    if (c[C_ERR] != 0) return fail(PYR_ERROR_CUDA, "build: device error flags 0x%x", c[C_ERR]);
    memset(out, 0, sizeof(*out));
    out->log2_size = g->log2_size;
    out->root = c[C_ROOT];
    out->nodes = (uint64_t)(uintptr_t)g->gb.nodes;
    out->node_words = c[C_WORDS];
    out->leaves = (uint64_t)(uintptr_t)g->gb.leaves;
    out->leaf_count = c[C_LEAVES];
    const uint32_t fin = c[C_FINAL] & 1;
    out->keys = (uint64_t)(uintptr_t)g->gb.kb[fin ^ 1];
    out->attrs = (uint64_t)(uintptr_t)(g->gb.packed ? g->gb.vb[0] : g->gb.vb[fin ^ 1]);
    out->voxel_count = c[C_U];
    memcpy(roots, c + C_TILE, (uint64_t)g->count * 4);
    memcpy(first_voxel, c + C_TILE + g->count, (uint64_t)g->count * 4);
    return PYR_OK;
}

PYR_API void pyrhip_built_free(PyrhipBuilt* b, void* stream) {
    if (!b) return;
    hipStream_t s = (hipStream_t)stream;
    pool_free((void*)(uintptr_t)b->nodes, s);
    pool_free((void*)(uintptr_t)b->leaves, s);
    pool_free((void*)(uintptr_t)b->keys, s);
    pool_free((void*)(uintptr_t)b->attrs, s);
    memset(b, 0, sizeof(*b));
}

PYR_API PyrResult pyrhip_decode_voxels(uint64_t keys, uint64_t attrs, uint32_t count, uint32_t log2_size,
                                       uint64_t out, void* stream) {
    diag_clear();
    if (count && (!keys || !attrs || !out)) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if (count)
        hipLaunchKernelGGL(k_decode, dim3(grid_for(count, kBlock * 4, 4096)), dim3(kBlock), 0, (hipStream_t)stream,
                           (const uint64_t*)(uintptr_t)keys, (const uint32_t*)(uintptr_t)attrs, count,
                           local_bits(log2_size), 0u, (int4*)(uintptr_t)out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "decode: %s", hipGetErrorString(e));
}

// Source voxels in ascending key order move to their coarse cells; the
// build's stable sort and last-record-wins unique then keep, per cell, the
// source voxel with the largest key.
PYR_API PyrResult pyrhip_downsample(uint64_t src_keys, uint64_t src_vals, uint32_t n_src, uint32_t log2_src,
                                    uint32_t shift, const PyrhipBuildTuning* tuning, void* stream, PyrhipBuilt* out) {
    diag_clear();
    if (!out || (n_src && (!src_keys || !src_vals))) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_src < 3 || log2_src > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if (shift == 0 || shift + 3 > log2_src)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "downsampling 2^%u by 2^%u is not possible (result at least 2^3)",
                    log2_src, shift);
    hipStream_t s = (hipStream_t)stream;
    void* vox = nullptr;
    if (pool_alloc(&vox, std::max<uint64_t>(n_src, 1) * 16, s) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_OUT_OF_MEMORY, "downsample input of %u voxels", n_src);
    }
    if (n_src)
        hipLaunchKernelGGL(k_decode, dim3(grid_for(n_src, kBlock * 4, 4096)), dim3(kBlock), 0, s,
                           (const uint64_t*)(uintptr_t)src_keys, (const uint32_t*)(uintptr_t)src_vals, n_src,
                           local_bits(log2_src), shift, (int4*)vox);
    hipError_t e = hipGetLastError();
    PyrResult r = e == hipSuccess ? pyrhip_build(log2_src - shift, 0, 0, 0, (uint64_t)(uintptr_t)vox, n_src, tuning,
                                                 stream, out)
                                  : fail(PYR_ERROR_CUDA, "downsample: %s", hipGetErrorString(e));
    pool_free(vox, s);
    return r;
}

PYR_API PyrResult pyrhip_copy_probe(uint64_t src, uint64_t dst, uint64_t bytes, void* stream) {
    diag_clear();
    if (!src || !dst || (bytes & 15) || ((src | dst) & 15))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "copy probe needs 16-byte aligned buffers and sizes");
    const uint64_t n = bytes / 16;
    const uint64_t blocks = (n + kProbeChunk - 1) / kProbeChunk;
    if (blocks > 0x7FFFFFFFull) return fail(PYR_ERROR_INVALID_ARGUMENT, "copy probe size");
    if (n)
        hipLaunchKernelGGL(k_copy_probe, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                           (const u32x4*)(uintptr_t)src, (u32x4*)(uintptr_t)dst, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "copy probe: %s", hipGetErrorString(e));
}

PYR_API void pyrhip_set_timing(int enable) { g_timing = enable != 0; }


PYR_API void pyrhip_last_build_times(PyrhipBuildTimes* out) {
    if (!out) return;
    std::lock_guard<std::mutex> lk(g_times_mutex);
    *out = g_times;
}

}  // extern "C"
