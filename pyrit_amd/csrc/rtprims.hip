// rtprims.hip - axis-aligned box primitives of a finished DAG: every
// occupied subtree of edge 2^k becomes one primitive {node, attribute rank
// of its first voxel, cell} with the tight box of its voxels (the reference
// extracts them on the CPU in src/rt_prims.zig:47-137 and emits the same
// list during its GPU build, src/device/gbuild.zig:574-605, 251-257, for its
// hardware ray-tracing path).
//
// The reference's extraction is a depth-first recursion with a hash memo of
// boxes per (level, node). Here it is three level-synchronous phases, each a
// few launches over arrays:
//   1. positions: the subtrees of level k in depth-first order, by expanding
//      a frontier {node, rank, origin} from the root one level at a time;
//      children go to the offsets of an exclusive scan of the per-item child
//      counts, so the order stays depth-first (= Morton order of the cells)
//      and ranks follow from the count words of earlier siblings;
//   2. distinct nodes: the set of nodes reachable at each level from k down
//      to 3 (a node shared by many cells is there once), one open-addressing
//      table per level;
//   3. boxes bottom-up over those sets: level 3 from the 4^3 brick masks
//      (per-axis plane tests, no loop over voxels), each higher level from
//      its children's boxes found in the level below's table.
// The primitives are then the positions whose box is not empty, compacted by
// one more scan. The host reads back each level's size to size the next
// level's arrays (one synchronisation per level); work and memory are linear
// in the positions and distinct nodes, never in the voxels.
//
// Every node and leaf reference is checked against the DAG's word and leaf
// counts before it is read; a bad reference sets an error word and the
// call fails with PYR_ERROR_INVALID_ARGUMENT.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "internal.h"

namespace pyr {
namespace {

constexpr int kItems = 4;
constexpr uint32_t kTile = kBlock * kItems;  // items per block of the scan kernels
constexpr uint32_t kCoord = 0x1FFFFFu;       // 21 bits per axis (edges up to 2^20)
constexpr uint32_t kErrRef = 1u;             // a node or leaf reference out of range
constexpr uint32_t kErrSet = 2u;             // a table lookup missed (inconsistent DAG)
constexpr uint64_t kMaxPositions = 1ull << 30;

struct Frontier {
    uint32_t* node;
    uint32_t* rank;
    uint64_t* origin;  // x | y << 21 | z << 42, voxel units
};

struct DagRef {
    const uint32_t* nodes;
    const uint64_t* leaves;
    uint32_t node_words, leaf_count;
};

// Open-addressing set of node words of one level: key = node + 1 (0 = empty),
// val = index of the node in the level's list.
struct NodeSet {
    uint32_t* keys;
    uint32_t* vals;
    uint32_t mask;     // capacity - 1, capacity a power of two
    uint32_t* list;    // distinct nodes
    uint32_t* count;   // device counter of list entries
    uint32_t limit;    // list capacity
    uint64_t* boxes;   // [2 i] = lo, [2 i + 1] = hi, packed like origins, hi exclusive
};

__device__ __forceinline__ uint64_t pack3(uint32_t x, uint32_t y, uint32_t z) {
    return (uint64_t)x | (uint64_t)y << 21 | (uint64_t)z << 42;
}
__device__ __forceinline__ uint32_t axis(uint64_t p, int a) { return (uint32_t)(p >> (21 * a)) & kCoord; }

__device__ __forceinline__ uint32_t set_hash(uint32_t key) {
    const uint32_t h = key * 0x9E3779B1u;
    return h ^ (h >> 16);
}

// child count of a node, 0 (and an error) when its words run past the DAG
__device__ __forceinline__ uint32_t child_count(const DagRef& g, uint32_t node, uint32_t* err) {
    if (node >= g.node_words) {
        atomicOr(err, kErrRef);
        return 0;
    }
    const uint32_t c = __popc(g.nodes[node] & 0xFFu);
    if ((uint64_t)node + 2 + c > g.node_words) {
        atomicOr(err, kErrRef);
        return 0;
    }
    return c;
}

__device__ __forceinline__ uint32_t set_find(const NodeSet& s, uint32_t node) {
    const uint32_t key = node + 1;
    uint32_t h = set_hash(key) & s.mask;
    for (uint32_t probe = 0; probe <= s.mask; probe++) {
        const uint32_t cur = s.keys[h];
        if (cur == key) return s.vals[h];
        if (cur == 0) break;
        h = (h + 1) & s.mask;
    }
    return 0xFFFFFFFFu;
}

__device__ __forceinline__ void set_insert(const NodeSet& s, uint32_t node, uint32_t* err) {
    const uint32_t key = node + 1;
    uint32_t h = set_hash(key) & s.mask;
    for (uint32_t probe = 0; probe <= s.mask; probe++) {
        uint32_t cur = __hip_atomic_load(&s.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == 0) {
            cur = atomicCAS(&s.keys[h], 0u, key);
            if (cur == 0) {
                const uint32_t idx = atomicAdd(s.count, 1u);
                if (idx >= s.limit) {  // more distinct nodes than the DAG's words allow
                    atomicOr(err, kErrRef);
                    return;
                }
                s.list[idx] = node;
                s.vals[h] = idx;
                return;
            }
        }
        if (cur == key) return;
        h = (h + 1) & s.mask;
    }
    atomicOr(err, kErrSet);  // table full: cannot happen at load <= 1/2
}

// item i of lane l in wave w of a tile: wave-striped, as tile_scan expects
__device__ __forceinline__ uint64_t item_index(uint32_t tile, int i) {
    return (uint64_t)tile * kTile + (uint64_t)(threadIdx.x / kWave) * kWave * kItems + (uint64_t)i * kWave + lane_id();
}

// block total of per-item counts
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s_wave) {
    v = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = v;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) t += s_wave[w];
    return t;
}

// ---- phase 1: frontier expansion ------------------------------------------

__global__ void __launch_bounds__(kBlock) k_expand_count(DagRef g, const uint32_t* node, uint32_t n, uint32_t* sums,
                                                         uint32_t* err) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t j = item_index(blockIdx.x, i);
        if (j < n) v += child_count(g, node[j], err);
    }
    const uint32_t t = block_sum(v, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// exclusive scan of the per-block sums (one block): offsets[b], total
__global__ void __launch_bounds__(kBlock) k_scan_sums(const uint32_t* sums, uint32_t nb, uint64_t* offsets,
                                                      uint64_t* total) {
    __shared__ uint64_t s_part[kBlock];
    const uint32_t per = (nb + kBlock - 1) / kBlock;
    const uint32_t lo = min(nb, threadIdx.x * per), hi = min(nb, lo + per);
    uint64_t own = 0;
    for (uint32_t b = lo; b < hi; b++) own += sums[b];
    s_part[threadIdx.x] = own;
    __syncthreads();
    for (uint32_t d = 1; d < kBlock; d *= 2) {  // inclusive Hillis-Steele scan
        const uint64_t add = threadIdx.x >= d ? s_part[threadIdx.x - d] : 0;
        __syncthreads();
        s_part[threadIdx.x] += add;
        __syncthreads();
    }
    uint64_t run = s_part[threadIdx.x] - own;
    for (uint32_t b = lo; b < hi; b++) {
        offsets[b] = run;
        run += sums[b];
    }
    if (threadIdx.x == kBlock - 1) *total = s_part[kBlock - 1];
}

__global__ void __launch_bounds__(kBlock) k_expand_write(DagRef g, Frontier in, uint32_t n, const uint64_t* offsets,
                                                         Frontier out, uint32_t child_log2, uint32_t* err) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    uint32_t cnt[kItems], excl[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t j = item_index(blockIdx.x, i);
        cnt[i] = j < n ? child_count(g, in.node[j], err) : 0u;
    }
    tile_scan<kItems>(cnt, excl, s_wave);
    const uint64_t base = offsets[blockIdx.x];
    const uint32_t child = 1u << child_log2;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        if (cnt[i] == 0) continue;
        const uint64_t j = item_index(blockIdx.x, i);
        const uint32_t node = in.node[j], mask = g.nodes[node] & 0xFFu;
        const uint64_t org = in.origin[j];
        uint32_t rank = in.rank[j];
        uint64_t pos = base + excl[i];
        uint32_t slot = 0;
        for (uint32_t idx = 0; idx < 8; idx++) {
            if (!(mask >> idx & 1u)) continue;
            const uint32_t ref = g.nodes[node + 2 + slot++];
            out.node[pos] = ref;
            out.rank[pos] = rank;
            out.origin[pos] = org + pack3((idx & 1u) * child, (idx >> 1 & 1u) * child, (idx >> 2) * child);
            pos++;
            if (ref + 1ull < g.node_words) {
                rank += g.nodes[ref + 1];  // wraps like the reference's u32 rank
            } else {
                atomicOr(err, kErrRef);
            }
        }
    }
}

// ---- phase 2: distinct nodes per level -------------------------------------

// Positions and children repeat a few nodes many times over (flat or
// periodic regions), so a block first merges its nodes in an LDS table of
// twice as many slots: only the first lane of each distinct node inserts it
// into the level's table, whose hot slots then see one access per block
// instead of one per occurrence.
constexpr uint32_t kLdsSlots = 2 * kTile;  // a block merges at most kTile nodes

__device__ __forceinline__ bool first_in_block(uint32_t* s_key, uint32_t key) {
    uint32_t h = set_hash(key) & (kLdsSlots - 1);
    for (uint32_t probe = 0; probe < kLdsSlots; probe++) {  // the table is never more than half full
        const uint32_t cur = atomicCAS(&s_key[h], 0u, key);
        if (cur == 0) return true;
        if (cur == key) return false;
        h = (h + 1) & (kLdsSlots - 1);
    }
    return true;
}

__global__ void __launch_bounds__(kBlock) k_set_items(const uint32_t* node, uint32_t n, NodeSet s, uint32_t* err) {
    __shared__ uint32_t s_key[kLdsSlots];
    for (uint32_t i = threadIdx.x; i < kLdsSlots; i += kBlock) s_key[i] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kItems; q++) {
        const uint64_t j = (uint64_t)blockIdx.x * kTile + (uint64_t)q * kBlock + threadIdx.x;
        if (j < n && first_in_block(s_key, node[j] + 1)) set_insert(s, node[j], err);
    }
}

// children of the parent level's distinct nodes (all nodes: parent level > 3)
// (128 parents per block: at most 1024 children, as many as k_set_items merges)
constexpr uint32_t kParentsPerBlock = kTile / 8;

__global__ void __launch_bounds__(kBlock) k_set_children(DagRef g, NodeSet parent, uint32_t n, NodeSet s,
                                                         uint32_t* err) {
    __shared__ uint32_t s_key[kLdsSlots];
    for (uint32_t i = threadIdx.x; i < kLdsSlots; i += kBlock) s_key[i] = 0;
    __syncthreads();
    // two lanes per parent, four child slots each
    const uint64_t j = (uint64_t)blockIdx.x * kParentsPerBlock + threadIdx.x / 2;
    if (j >= n) return;
    const uint32_t node = parent.list[j];
    const uint32_t c = child_count(g, node, err);
    for (uint32_t q = (threadIdx.x & 1) * 4; q < min(c, (threadIdx.x & 1) * 4 + 4); q++) {
        const uint32_t child = g.nodes[node + 2 + q];
        if (first_in_block(s_key, child + 1)) set_insert(s, child, err);
    }
}

// ---- phase 3: boxes bottom-up ----------------------------------------------

// occupancy of the four planes along each axis of a 4^3 brick (bit x + 4y + 16z)
__device__ __forceinline__ void brick_box(uint64_t m, uint32_t lo[3], uint32_t hi[3]) {
    uint32_t occ[3] = {0, 0, 0};
#pragma unroll
    for (uint32_t p = 0; p < 4; p++) {
        occ[0] |= (m & (0x1111111111111111ull << p)) ? 1u << p : 0u;
        occ[1] |= (m & (0x000F000F000F000Full << (4 * p))) ? 1u << p : 0u;
        occ[2] |= (m & (0xFFFFull << (16 * p))) ? 1u << p : 0u;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = occ[a] ? (uint32_t)__builtin_ctz(occ[a]) : kCoord;
        hi[a] = occ[a] ? 32u - (uint32_t)__builtin_clz(occ[a]) : 0u;
    }
}

__global__ void __launch_bounds__(kBlock) k_boxes(DagRef g, NodeSet s, uint32_t n, NodeSet below, uint32_t level,
                                                  uint32_t* err) {
    const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t node = s.list[j];
    const uint32_t c = child_count(g, node, err);
    const uint32_t mask = c ? g.nodes[node] & 0xFFu : 0u, child = 1u << (level - 1);
    uint32_t lo[3] = {kCoord, kCoord, kCoord}, hi[3] = {0, 0, 0};
    uint32_t slot = 0;
    for (uint32_t idx = 0; idx < 8; idx++) {
        if (!(mask >> idx & 1u)) continue;
        const uint32_t ref = g.nodes[node + 2 + slot++];
        uint32_t clo[3], chi[3];
        if (level == 3) {
            if (ref >= g.leaf_count) {
                atomicOr(err, kErrRef);
                continue;
            }
            brick_box(g.leaves[ref], clo, chi);
        } else {
            const uint32_t q = set_find(below, ref);
            if (q == 0xFFFFFFFFu) {
                atomicOr(err, kErrSet);
                continue;
            }
            const uint64_t bl = below.boxes[2ull * q], bh = below.boxes[2ull * q + 1];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                clo[a] = axis(bl, a);
                chi[a] = axis(bh, a);
            }
        }
        if (clo[0] >= chi[0]) continue;  // empty child
        const uint32_t off[3] = {(idx & 1u) * child, (idx >> 1 & 1u) * child, (idx >> 2) * child};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = min(lo[a], clo[a] + off[a]);
            hi[a] = max(hi[a], chi[a] + off[a]);
        }
    }
    s.boxes[2 * j] = pack3(lo[0], lo[1], lo[2]);
    s.boxes[2 * j + 1] = pack3(hi[0], hi[1], hi[2]);
}

// ---- primitives: positions with a non-empty box, compacted in order ---------

__device__ __forceinline__ bool prim_box(const NodeSet& s, uint32_t node, uint64_t* bl, uint64_t* bh, uint32_t* err) {
    const uint32_t q = set_find(s, node);
    if (q == 0xFFFFFFFFu) {
        atomicOr(err, kErrSet);
        return false;
    }
    *bl = s.boxes[2ull * q];
    *bh = s.boxes[2ull * q + 1];
    return axis(*bl, 0) < axis(*bh, 0);
}

__global__ void __launch_bounds__(kBlock) k_prim_count(Frontier in, uint32_t n, NodeSet s, uint32_t* sums,
                                                       uint32_t* err) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t j = item_index(blockIdx.x, i);
        uint64_t bl, bh;
        if (j < n && prim_box(s, in.node[j], &bl, &bh, err)) v++;
    }
    const uint32_t t = block_sum(v, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock) k_prim_write(Frontier in, uint32_t n, NodeSet s, const uint64_t* offsets,
                                                       uint32_t k, PyrhipRtPrim* prims, float* aabbs, uint32_t* err) {
    __shared__ uint32_t s_wave[kBlock / kWave];
    uint32_t keep[kItems], excl[kItems];
    uint64_t bl[kItems], bh[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t j = item_index(blockIdx.x, i);
        keep[i] = j < n && prim_box(s, in.node[j], &bl[i], &bh[i], err) ? 1u : 0u;
    }
    tile_scan<kItems>(keep, excl, s_wave);
    const uint64_t base = offsets[blockIdx.x];
    const float eps = 1.0f / 256.0f;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        if (!keep[i]) continue;
        const uint64_t j = item_index(blockIdx.x, i), p = base + excl[i];
        const uint64_t org = in.origin[j];
        PyrhipRtPrim r;
        r.node = in.node[j];
        r.attr_base = in.rank[j];
        r.cell = (uint64_t)(axis(org, 0) >> k) | (uint64_t)(axis(org, 1) >> k) << 21 |
                 (uint64_t)(axis(org, 2) >> k) << 42;
        prims[p] = r;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            aabbs[6 * p + a] = (float)(axis(org, a) + axis(bl[i], a)) - eps;
            aabbs[6 * p + 3 + a] = (float)(axis(org, a) + axis(bh[i], a)) + eps;
        }
    }
}

// ---- host ------------------------------------------------------------------

// device allocations of one call, returned to the cache when it ends
struct Scratch {
    hipStream_t s;
    std::vector<void*> blocks;
    explicit Scratch(hipStream_t st) : s(st) {}
    ~Scratch() {
        for (void* p : blocks) pool_free(p, s);
    }
    template <class T>
    T* get(uint64_t count) {
        void* p = nullptr;
        if (pool_alloc(&p, std::max<uint64_t>(count * sizeof(T), 16), s) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        blocks.push_back(p);
        return (T*)p;
    }
};

uint32_t tiles_of(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }
uint32_t blocks_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace
}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_rt_prims(const PyrhipDag* dag, uint32_t node_words, uint32_t leaf_count,
                                  uint32_t requested_log2, uint64_t prims, uint64_t aabbs, uint32_t capacity,
                                  void* stream, uint32_t* out_count, uint32_t* out_log2) {
    diag_clear();
    if (!dag || !out_count || !dag->nodes || (leaf_count && !dag->leaves))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->log2_size < 3 || dag->log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size out of range");
    if (dag->root >= node_words) return fail(PYR_ERROR_INVALID_ARGUMENT, "root outside the node words");
    if ((prims != 0) != (aabbs != 0)) return fail(PYR_ERROR_INVALID_ARGUMENT, "prims and aabbs go together");
    const uint32_t L = dag->log2_size;
    const uint32_t k = std::min(std::max(requested_log2, 3u), L);
    if (out_log2) *out_log2 = k;
    *out_count = 0;
    hipStream_t s = (hipStream_t)stream;
    Scratch sc(s);
    const DagRef g{(const uint32_t*)(uintptr_t)dag->nodes, (const uint64_t*)(uintptr_t)dag->leaves, node_words,
                   leaf_count};
    // [0] error word, [2..3] u64 total of the last scan
    uint32_t* ctl = sc.get<uint32_t>(4);
    if (!ctl) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims control words");
    uint32_t* err = ctl;
    uint64_t* total = (uint64_t*)(ctl + 2);
    struct Host {
        uint32_t err, pad;
        uint64_t total;
    } h{};
    auto sync_read = [&](const char* what) -> PyrResult {
        if (hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            (void)hipGetLastError();
            return fail(PYR_ERROR_CUDA, "rt prims: %s", what);
        }
        if (h.err & kErrRef) return fail(PYR_ERROR_INVALID_ARGUMENT, "rt prims: a node or leaf reference out of range");
        if (h.err) return fail(PYR_ERROR_INVALID_ARGUMENT, "rt prims: inconsistent DAG (a child is missing)");
        return PYR_OK;
    };
    if (hipMemsetAsync(ctl, 0, 16, s) != hipSuccess) return fail(PYR_ERROR_CUDA, "rt prims clear");

    // phase 1: positions of level k, depth-first
    auto frontier = [&](uint64_t n, Frontier* f) {
        f->node = sc.get<uint32_t>(n);
        f->rank = sc.get<uint32_t>(n);
        f->origin = sc.get<uint64_t>(n);
        return f->node && f->rank && f->origin;
    };
    Frontier cur{};
    if (!frontier(1, &cur)) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims frontier");
    const uint32_t zero = 0;
    const uint64_t zero64 = 0;
    if (hipMemcpyAsync(cur.node, &dag->root, 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(cur.rank, &zero, 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(cur.origin, &zero64, 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(PYR_ERROR_CUDA, "rt prims upload");
    uint64_t n = 1;
    uint32_t* sums = nullptr;
    uint64_t* offsets = nullptr;
    uint32_t sums_cap = 0;
    auto scan_space = [&](uint32_t nb) {
        if (nb <= sums_cap) return true;
        sums = sc.get<uint32_t>(nb);
        offsets = sc.get<uint64_t>(nb);
        sums_cap = nb;
        return sums && offsets;
    };
    for (uint32_t level = L; level > k && n; level--) {
        const uint32_t nb = tiles_of(n);
        if (!scan_space(nb)) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims scan");
        hipLaunchKernelGGL(k_expand_count, dim3(nb), dim3(kBlock), 0, s, g, cur.node, (uint32_t)n, sums, err);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, sums, nb, offsets, total);
        if (PyrResult r = sync_read("expansion"); r != PYR_OK) return r;
        if (h.total > kMaxPositions)
            return fail(PYR_ERROR_CAPACITY, "rt prims: %llu subtrees of edge 2^%u, more than 2^30",
                        (unsigned long long)h.total, level - 1);
        Frontier next{};
        if (h.total && !frontier(h.total, &next)) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims frontier");
        if (h.total)
            hipLaunchKernelGGL(k_expand_write, dim3(nb), dim3(kBlock), 0, s, g, cur, (uint32_t)n, offsets, next,
                               level - 1, err);
        cur = next;
        n = h.total;
    }
    if (n == 0) return hipStreamSynchronize(s) == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "rt prims");

    // phase 2: distinct nodes of levels k .. 3
    std::vector<NodeSet> sets(k + 1);
    std::vector<uint32_t> size(k + 1, 0);
    auto make_set = [&](uint64_t bound, NodeSet* st) {
        uint64_t cap = 64;
        while (cap < 2 * bound) cap *= 2;
        st->keys = sc.get<uint32_t>(2 * cap);  // keys then vals
        st->list = sc.get<uint32_t>(bound + 1);
        st->boxes = sc.get<uint64_t>(2 * bound);
        if (!st->keys || !st->list || !st->boxes) return false;
        st->vals = st->keys + cap;
        st->mask = (uint32_t)(cap - 1);
        st->limit = (uint32_t)bound;
        st->count = st->list + bound;  // counter after the list
        return hipMemsetAsync(st->keys, 0, cap * 4, s) == hipSuccess &&
               hipMemsetAsync(st->count, 0, 4, s) == hipSuccess;
    };
    const uint64_t node_bound = node_words / 3 + 1;  // every node holds at least 3 words
    if (!make_set(std::min<uint64_t>(n, node_bound), &sets[k])) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims set");
    hipLaunchKernelGGL(k_set_items, dim3(tiles_of(n)), dim3(kBlock), 0, s, cur.node, (uint32_t)n, sets[k], err);
    for (uint32_t level = k;; level--) {
        uint32_t c = 0;
        if (hipMemcpyAsync(&c, sets[level].count, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
            return fail(PYR_ERROR_CUDA, "rt prims set size");
        if (PyrResult r = sync_read("node sets"); r != PYR_OK) return r;
        size[level] = c;
        if (level == 3) break;
        if (!make_set(std::min<uint64_t>(8ull * c, node_bound), &sets[level - 1]))
            return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims set");
        if (c)
            hipLaunchKernelGGL(k_set_children, dim3((c + kParentsPerBlock - 1) / kParentsPerBlock), dim3(kBlock), 0,
                               s, g, sets[level], c, sets[level - 1], err);
    }

    // phase 3: boxes, level 3 first
    for (uint32_t level = 3; level <= k; level++)
        if (size[level])
            hipLaunchKernelGGL(k_boxes, dim3(blocks_of(size[level])), dim3(kBlock), 0, s, g, sets[level],
                               size[level], level > 3 ? sets[level - 1] : sets[level], level, err);

    // primitives: positions with a non-empty box
    const uint32_t nb = tiles_of(n);
    if (!scan_space(nb)) return fail(PYR_ERROR_OUT_OF_MEMORY, "rt prims scan");
    hipLaunchKernelGGL(k_prim_count, dim3(nb), dim3(kBlock), 0, s, cur, (uint32_t)n, sets[k], sums, err);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, sums, nb, offsets, total);
    if (PyrResult r = sync_read("primitive count"); r != PYR_OK) return r;
    *out_count = (uint32_t)h.total;
    if (!prims) return PYR_OK;
    if (h.total > capacity)
        return fail(PYR_ERROR_CAPACITY, "rt prims: %llu primitives, capacity %u", (unsigned long long)h.total,
                    capacity);
    hipLaunchKernelGGL(k_prim_write, dim3(nb), dim3(kBlock), 0, s, cur, (uint32_t)n, sets[k], offsets, k,
                       (PyrhipRtPrim*)(uintptr_t)prims, (float*)(uintptr_t)aabbs, err);
    return sync_read("primitive write");
}

}  // extern "C"
