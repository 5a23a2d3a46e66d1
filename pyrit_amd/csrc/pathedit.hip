// pathedit.hip - voxel edits applied to a finished DAG without a rebuild
// (the capability of the reference's pyr_k_path_edit, src/device/pathedit.zig,
// driven by context.zig:1093+ and world.zig:394-460).
//
// Design. A job (one chunk) is a list of edits sorted by the Morton key of
// their brick, in submission order inside a brick. Bricks are independent:
// each edited brick's new mask is its old mask with its own edits applied in
// order, so one wave handles one job and its lanes work on different bricks:
//
//   k_path_walk   one wave per job. Lanes walk root -> brick for different
//                 bricks at once (independent chains of loads), recording the
//                 old node at every level, the old brick and the voxels that
//                 precede it in DFS order; apply the brick's edits; a wave
//                 prefix sum over the bricks gives each brick's rank shift.
//                 Then the touched nodes are rewritten bottom-up, one level
//                 at a time: a node is touched when an edited brick lies
//                 below it (a run of equal key prefixes among the sorted
//                 bricks), its children are found by binary search in its
//                 run, and it is written again only when a child reference
//                 changed. New leaves and nodes take their arena slots from
//                 wave prefix sums of their sizes, so every lane stays busy.
//   k_path_attrs  the new attribute array, 256 ranks per block over all jobs:
//                 a rank finds its edited brick by binary search; ranks
//                 between edited bricks copy the old array shifted by the
//                 voxels inserted or removed before them.
//
// The copies go to an arena above every batch in the pools, so the chunk's
// ordinary relative references reach them and neither the node format nor
// the traversal changes. The old words stay valid until the caller releases
// them.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

constexpr uint32_t kNoRef = 0xFFFFFFFFu;  // absent child / node / leaf
constexpr uint32_t kMaxLevels = 18;       // node levels of a 2^20 grid
constexpr int kAttrBlock = 256;

// Per-brick state of all jobs, indexed by the job's edit_first + brick.
struct BrickArrays {
    uint64_t* key;        // brick Morton key
    uint64_t* old_mask;   // 0: the brick was absent
    uint64_t* new_mask;
    uint32_t* first;      // the brick's first edit (job-relative)
    uint32_t* old_rank;   // voxels before the brick in the old DFS order
    uint32_t* new_rank;   // ... in the new order
    int32_t* shift;       // voxels inserted minus removed up to and including the brick
    uint32_t* ref[2];     // subtree reference after the edits (run heads only), per level parity
    uint32_t* path;       // [level - 2][slot]: old node whose children have edge 2^level
    uint32_t slots;       // entries per array (= edits of all jobs)
};

__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ inline uint64_t brick_key(const uint4& e) {
    return spread3(e.x >> 2) | (spread3(e.y >> 2) << 1) | (spread3(e.z >> 2) << 2);
}

__device__ inline uint32_t brick_bit(const uint4& e) { return (e.x & 3u) + 4u * (e.y & 3u) + 16u * (e.z & 3u); }

// inclusive prefix sum over the 64 lanes of the (single-wave) block
__device__ inline uint32_t wave_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// first index in [lo, hi) whose value f(i) >= target (f nondecreasing)
template <class F>
__device__ inline uint32_t lower_bound(uint32_t lo, uint32_t hi, uint64_t target, F f) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (f(mid) < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One job per block of one wave.
__global__ void __launch_bounds__(64) k_path_walk(PathEditJob* jobs, uint32_t njobs, const uint4* edits,
                                                   BrickArrays B, uint32_t* nodes_pool, uint64_t* leaves_pool) {
    __shared__ uint32_t s_bricks, s_status;
    if (blockIdx.x >= njobs) return;
    PathEditJob& job = jobs[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    const uint32_t ne = job.edit_count, base = job.edit_first, L = job.log2_size;
    const uint4* ed = edits + base;
    uint32_t* nodes = nodes_pool + job.node_base;
    uint64_t* leaves = leaves_pool + job.leaf_base;
    uint64_t* key = B.key + base;
    uint64_t* omask = B.old_mask + base;
    uint64_t* nmask = B.new_mask + base;
    uint32_t* first = B.first + base;
    uint32_t* orank = B.old_rank + base;
    uint32_t* nrank = B.new_rank + base;
    int32_t* shift = B.shift + base;

    // bricks: heads of runs of equal brick keys among the sorted edits
    uint32_t carry = 0;
    for (uint32_t e0 = 0; e0 < ne; e0 += 64) {
        const uint32_t e = e0 + lane;
        uint64_t k = 0;
        bool head = false;
        if (e < ne) {
            k = brick_key(ed[e]);
            head = e == 0 || brick_key(ed[e - 1]) != k;
        }
        const uint32_t inc = wave_scan(head ? 1u : 0u);
        if (head) {
            key[carry + inc - 1] = k;
            first[carry + inc - 1] = e;
        }
        carry += __shfl(inc, 63, 64);
    }
    const uint32_t nb = carry;
    __syncthreads();

    // walk root -> brick for each brick, apply its edits
    for (uint32_t b = lane; b < nb; b += 64) {
        const uint4 lead = ed[first[b]];
        const uint32_t v[3] = {lead.x, lead.y, lead.z};
        uint32_t node = job.root, rank = 0, leaf = kNoRef;
        for (uint32_t s = L - 1; s >= 2; s--) {
            B.path[(uint64_t)(s - 2) * B.slots + base + b] = node;
            if (node == kNoRef) continue;
            const uint32_t m = nodes[node] & 0xFFu;
            const uint32_t c = ((v[0] >> s) & 1u) | (((v[1] >> s) & 1u) << 1) | (((v[2] >> s) & 1u) << 2);
            const uint32_t before = (uint32_t)__popc(m & ((1u << c) - 1u));
            for (uint32_t k = 0; k < before; k++) {
                const uint32_t r = nodes[node + 2 + k];
                rank += s == 2 ? (uint32_t)__popcll(leaves[r]) : nodes[r + 1];
            }
            const uint32_t child = (m >> c) & 1u ? nodes[node + 2 + before] : kNoRef;
            if (s == 2) leaf = child;
            else node = child;
        }
        const uint64_t old = leaf == kNoRef ? 0ull : leaves[leaf];
        uint64_t m = old;
        const uint32_t end = b + 1 < nb ? first[b + 1] : ne;
        for (uint32_t e = first[b]; e < end; e++) {
            const uint4 q = ed[e];
            const uint64_t bit = 1ull << brick_bit(q);
            m = q.w ? (m | bit) : (m & ~bit);
        }
        omask[b] = old;
        nmask[b] = m;
        orank[b] = rank;
        // the brick's reference after the edits: the old leaf while its mask
        // is unchanged, else a new leaf (slot below) or none
        B.ref[0][base + b] = m == old ? leaf : kNoRef;
    }
    __syncthreads();

    // rank shifts and new leaves: prefix sums over the bricks
    int32_t sh = 0;
    uint32_t leaf_next = job.leaf_arena, status = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
        const uint32_t b = b0 + lane;
        int32_t d = 0;
        bool fresh = false;
        if (b < nb) {
            d = __popcll(nmask[b]) - __popcll(omask[b]);
            fresh = nmask[b] != omask[b] && nmask[b] != 0;
        }
        const int32_t inc = (int32_t)wave_scan((uint32_t)d);  // two's complement sums
        const uint32_t slot = wave_scan(fresh ? 1u : 0u);
        if (b < nb) {
            shift[b] = sh + inc;
            nrank[b] = orank[b] + (uint32_t)(sh + inc - d);
            if (fresh) {
                const uint32_t at = leaf_next + slot - 1;
                if (at < job.leaf_end) {
                    leaves_pool[at] = nmask[b];
                    B.ref[0][base + b] = at - job.leaf_base;
                }
            }
        }
        sh += __shfl(inc, 63, 64);
        leaf_next += __shfl(slot, 63, 64);
    }
    if (leaf_next > job.leaf_end) status |= kPathOverflow;
    __syncthreads();

    // touched nodes, bottom-up: level s holds nodes whose children have edge
    // 2^s; a node is identified by key >> 3 (s - 1) and its child digit by
    // (key >> 3 (s - 2)) & 7 (children of level-2 nodes are the bricks)
    uint32_t node_next = job.node_arena;
    for (uint32_t s = 2; s < L && !status; s++) {
        const uint32_t up = 3 * (s - 1), dn = 3 * (s - 2);
        const uint32_t* ref_in = B.ref[s & 1] + base;   // children (level s - 1 heads)
        uint32_t* ref_out = B.ref[(s & 1) ^ 1] + base;  // this level's heads
        const uint32_t* path = B.path + (uint64_t)(s - 2) * B.slots + base;
        // pass 0: sizes of the nodes to write; pass 1: write them
        uint32_t total = 0;
        for (int pass = 0; pass < 2; pass++) {
            uint32_t cursor = node_next;
            for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
                const uint32_t h = b0 + lane;
                uint32_t words = 0, new_ref = kNoRef;
                uint32_t mask = 0, cref[8], count = 0, old = kNoRef;
                const bool head = h < nb && (h == 0 || (key[h] >> up) != (key[h - 1] >> up));
                if (head) {
                    const uint64_t id = key[h] >> up;
                    const uint32_t end = lower_bound(h + 1, nb, (id + 1) << up, [&](uint32_t i) { return key[i]; });
                    old = path[h];
                    const uint32_t om = old == kNoRef ? 0u : nodes[old] & 0xFFu;
                    bool changed = false;
                    uint32_t slot = 0, c = h;
                    for (uint32_t k = 0; k < 8; k++) {
                        const uint32_t oref = (om >> k) & 1u ? nodes[old + 2 + slot] : kNoRef;
                        slot += (om >> k) & 1u;
                        // the run's first brick with child digit > k
                        const uint32_t next = lower_bound(c, end, ((id << 3) + k + 1) << dn,
                                                          [&](uint32_t i) { return key[i]; });
                        uint32_t r = oref;
                        if (next > c) {  // edited bricks below child k: its head is c
                            r = ref_in[c];
                            changed |= r != oref;
                        }
                        c = next;
                        cref[k] = r;
                        if (r != kNoRef) mask |= 1u << k;
                    }
                    if (!changed) {
                        new_ref = old;
                    } else if (mask) {
                        words = 2 + (uint32_t)__popc(mask);
                        count = (old == kNoRef ? 0u : nodes[old + 1]) +
                                (uint32_t)(shift[end - 1] - (h ? shift[h - 1] : 0));
                    }
                }
                const uint32_t at = wave_scan(words);
                const uint32_t dst = cursor + at - words;
                if (pass == 1 && head) {
                    if (words) {
                        uint32_t* w = nodes_pool + dst;
                        w[0] = mask;
                        w[1] = count;
                        uint32_t n = 2;
                        for (uint32_t k = 0; k < 8; k++)
                            if (cref[k] != kNoRef) w[n++] = cref[k];
                        new_ref = dst - job.node_base;
                    }
                    ref_out[h] = new_ref;
                }
                cursor += __shfl(at, 63, 64);
            }
            total = cursor - node_next;
            if (pass == 0 && node_next + total > job.node_end) {
                status |= kPathOverflow;
                break;
            }
        }
        node_next += total;
        __syncthreads();
    }
    if (lane == 0) {
        s_bricks = nb;
        s_status = status;
    }
    __syncthreads();
    if (lane == 0) {
        const uint32_t root = s_status ? kNoRef : B.ref[L & 1][base];
        job.bricks = s_bricks;
        job.new_count = job.attr_count + (uint32_t)(nb ? shift[nb - 1] : 0);
        job.new_root = root;
        job.status = s_status ? s_status : (root == kNoRef ? kPathEmpty : 0u);
    }
}

// The new attribute array: block t covers ranks [tile_first[t], +256) of job
// tile_job[t]. A rank inside an edited brick takes the value of the last edit
// of its voxel, or its old value; other ranks copy the old array shifted by
// the voxels inserted and removed before them.
__global__ void __launch_bounds__(kAttrBlock) k_path_attrs(const PathEditJob* jobs, const uint32_t* tile_job,
                                                            const uint32_t* tile_first, const uint4* edits,
                                                            BrickArrays B, const uint32_t* src_pool,
                                                            uint32_t* dst_pool) {
    const PathEditJob& job = jobs[tile_job[blockIdx.x]];
    const uint32_t i = tile_first[blockIdx.x] + threadIdx.x;
    if (job.status || i >= job.new_count) return;
    const uint32_t base = job.edit_first, nb = job.bricks;
    const uint32_t* nrank = B.new_rank + base;
    // last edited brick starting at or before rank i
    const uint32_t b1 = lower_bound(0, nb, (uint64_t)i + 1, [&](uint32_t k) { return nrank[k]; });
    uint32_t old_rank = i;
    bool have = false;
    uint32_t val = 0;
    if (b1) {
        const uint32_t b = b1 - 1;
        const uint64_t m = B.new_mask[base + b];
        const uint32_t j = i - nrank[b];
        if (j < (uint32_t)__popcll(m)) {
            // the j-th set bit of the new mask
            uint64_t r = m;
            for (uint32_t k = 0; k < j; k++) r &= r - 1;
            const uint32_t bit = (uint32_t)__ffsll((unsigned long long)r) - 1;
            const uint32_t e0 = B.first[base + b];
            const uint32_t e1 = b + 1 < nb ? B.first[base + b + 1] : job.edit_count;
            for (uint32_t e = e1; e-- > e0;) {
                const uint4 q = edits[base + e];
                if (brick_bit(q) == bit) {  // the last edit of this voxel set it
                    val = q.w;
                    have = true;
                    break;
                }
            }
            old_rank = B.old_rank[base + b] + (uint32_t)__popcll(B.old_mask[base + b] & ((1ull << bit) - 1));
        } else {
            old_rank = i - (uint32_t)B.shift[base + b];
        }
    }
    if (!have) {
        const uint32_t* src = src_pool + job.attr_src;
        val = job.palette_src != kNoRef
                  ? src_pool[job.palette_src + ((src[old_rank >> 3] >> ((old_rank & 7u) * 4u)) & 15u)]
                  : src[old_rank];
    }
    dst_pool[job.attr_dst + i] = val;
}

}  // namespace

// The edits of every job sorted by brick Morton key, in their order inside a
// brick (the walk and the attribute kernel take them so).
void sort_path_edits(const uint32_t* edits, uint32_t count, std::vector<uint32_t>* out) {
    auto spread = [](uint32_t v) {
        uint64_t x = 0;
        for (uint32_t b = 0; b < 21; b++) x |= (uint64_t)((v >> b) & 1u) << (3 * b);
        return x;
    };
    std::vector<std::pair<uint64_t, uint32_t>> order(count);
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t* e = edits + (size_t)i * 4;
        order[i] = {spread(e[0] >> 2) | (spread(e[1] >> 2) << 1) | (spread(e[2] >> 2) << 2), i};
    }
    std::sort(order.begin(), order.end());  // (key, index): stable by construction
    const size_t at = out->size();
    out->resize(at + (size_t)count * 4);
    for (uint32_t i = 0; i < count; i++)
        std::copy(edits + (size_t)order[i].second * 4, edits + (size_t)order[i].second * 4 + 4,
                  out->data() + at + (size_t)i * 4);
}

// Arena words and leaves a job can need at most: one leaf per edited brick,
// and per level one node of at most 10 words per distinct key prefix.
void path_edit_capacity(const uint32_t* sorted, uint32_t count, uint32_t log2_size, uint64_t* node_words,
                        uint64_t* leaves) {
    auto key = [&](uint32_t i) {
        const uint32_t* e = sorted + (size_t)i * 4;
        uint64_t k = 0;
        for (uint32_t b = 0; b < 21; b++)
            k |= (uint64_t)(((e[0] >> 2) >> b) & 1u) << (3 * b) | (uint64_t)(((e[1] >> 2) >> b) & 1u) << (3 * b + 1) |
                 (uint64_t)(((e[2] >> 2) >> b) & 1u) << (3 * b + 2);
        return k;
    };
    std::vector<uint64_t> keys;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t k = key(i);
        if (keys.empty() || keys.back() != k) keys.push_back(k);
    }
    *leaves = keys.size();
    uint64_t words = 0;
    for (uint32_t s = 2; s < log2_size; s++) {
        const uint32_t up = 3 * (s - 1);
        uint64_t distinct = 0;
        for (size_t i = 0; i < keys.size(); i++) distinct += i == 0 || (keys[i] >> up) != (keys[i - 1] >> up);
        words += distinct * 10;
    }
    *node_words = words;
}

PyrResult run_path_edits(std::vector<PathEditJob>& jobs, const std::vector<uint32_t>& edits, uint32_t* nodes_pool,
                         uint64_t* leaves_pool, const uint32_t* src_pool, uint32_t* dst_pool, hipStream_t s) {
    if (jobs.empty()) return PYR_OK;
    const uint32_t slots = std::max<uint32_t>((uint32_t)(edits.size() / 4), 1);
    // attribute tiles: every rank a job can end with (old count + its edits)
    std::vector<uint32_t> tile_job, tile_first;
    for (uint32_t j = 0; j < jobs.size(); j++) {
        if (jobs[j].log2_size < 3 || jobs[j].log2_size - 2 > kMaxLevels)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "path edit: log2_size %u", jobs[j].log2_size);
        const uint64_t most = (uint64_t)jobs[j].attr_count + jobs[j].edit_count;
        for (uint64_t r = 0; r < most; r += kAttrBlock) {
            tile_job.push_back(j);
            tile_first.push_back((uint32_t)r);
        }
    }
    const uint32_t tiles = (uint32_t)tile_job.size();
    std::vector<void*> held;
    bool ok = true;
    auto get = [&](uint64_t bytes) -> void* {
        void* p = nullptr;
        if (pool_alloc(&p, std::max<uint64_t>(bytes, 256), s) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        held.push_back(p);
        return p;
    };
    PathEditJob* d_jobs = (PathEditJob*)get(jobs.size() * sizeof(PathEditJob));
    uint4* d_edits = (uint4*)get(edits.size() * 4);
    uint32_t* d_tiles = (uint32_t*)get((uint64_t)std::max<uint32_t>(tiles, 1) * 8);
    BrickArrays B{};
    B.slots = slots;
    B.key = (uint64_t*)get((uint64_t)slots * 8);
    B.old_mask = (uint64_t*)get((uint64_t)slots * 8);
    B.new_mask = (uint64_t*)get((uint64_t)slots * 8);
    B.first = (uint32_t*)get((uint64_t)slots * 4);
    B.old_rank = (uint32_t*)get((uint64_t)slots * 4);
    B.new_rank = (uint32_t*)get((uint64_t)slots * 4);
    B.shift = (int32_t*)get((uint64_t)slots * 4);
    B.ref[0] = (uint32_t*)get((uint64_t)slots * 4);
    B.ref[1] = (uint32_t*)get((uint64_t)slots * 4);
    B.path = (uint32_t*)get((uint64_t)slots * 4 * kMaxLevels);
    PyrResult r = PYR_OK;
    if (!ok) {
        (void)hipGetLastError();
        r = fail(PYR_ERROR_OUT_OF_MEMORY, "path edit scratch");
    }
    if (r == PYR_OK) {
        std::vector<uint32_t> tl(tile_job);
        tl.insert(tl.end(), tile_first.begin(), tile_first.end());
        if (hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(PathEditJob), hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            (!edits.empty() &&
             hipMemcpyAsync(d_edits, edits.data(), edits.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) ||
            (tiles && hipMemcpyAsync(d_tiles, tl.data(), tl.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess))
            r = fail(PYR_ERROR_CUDA, "path edit upload");
    }
    if (r == PYR_OK) {
        hipLaunchKernelGGL(k_path_walk, dim3((uint32_t)jobs.size()), dim3(64), 0, s, d_jobs, (uint32_t)jobs.size(),
                           d_edits, B, nodes_pool, leaves_pool);
        if (tiles)
            hipLaunchKernelGGL(k_path_attrs, dim3(tiles), dim3(kAttrBlock), 0, s, d_jobs, d_tiles, d_tiles + tiles,
                               d_edits, B, src_pool, dst_pool);
        // (the host copies above must finish before their buffers go away)
        if (hipMemcpyAsync(jobs.data(), d_jobs, jobs.size() * sizeof(PathEditJob), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            r = fail(PYR_ERROR_CUDA, "path edit");
    }
    if (r != PYR_OK) (void)hipGetLastError();
    for (void* p : held) pool_free(p, s);
    return r;
}

}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_path_edit(const PyrhipDag* dag, uint32_t voxel_count, uint32_t node_words,
                                   uint32_t node_capacity, uint32_t leaf_count, uint32_t leaf_capacity,
                                   const uint32_t* edits, uint32_t count, uint64_t attrs_out, void* stream,
                                   uint32_t* out_root, uint32_t* out_count, uint32_t* out_flags) {
    diag_clear();
    if (!dag || !out_root || !out_count || !out_flags || (count && !edits) || !dag->nodes || !dag->attrs ||
        !attrs_out)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->palette) return fail(PYR_ERROR_INVALID_ARGUMENT, "palette attributes: unpack them first");
    if (dag->log2_size < 3 || dag->log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size out of range");
    if (dag->root >= node_words || node_words > node_capacity || leaf_count > leaf_capacity)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "root, words or capacities inconsistent");
    const uint32_t n = 1u << dag->log2_size;
    for (uint32_t i = 0; i < count; i++)
        if (edits[i * 4] >= n || edits[i * 4 + 1] >= n || edits[i * 4 + 2] >= n)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "edit %u outside the DAG", i);
    if (count == 0) {
        *out_root = dag->root;
        *out_count = voxel_count;
        *out_flags = 0;
        return hipMemcpyAsync((void*)(uintptr_t)attrs_out, (const void*)(uintptr_t)dag->attrs,
                              (uint64_t)voxel_count * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess &&
                       hipStreamSynchronize((hipStream_t)stream) == hipSuccess
                   ? PYR_OK
                   : fail(PYR_ERROR_CUDA, "attribute copy");
    }
    std::vector<uint32_t> sorted;
    sort_path_edits(edits, count, &sorted);
    std::vector<PathEditJob> jobs(1);
    PathEditJob& j = jobs[0];
    j.root = dag->root;
    j.log2_size = dag->log2_size;
    j.node_arena = node_words;
    j.node_end = node_capacity;
    j.leaf_arena = leaf_count;
    j.leaf_end = leaf_capacity;
    j.attr_count = voxel_count;
    j.edit_count = count;
    PyrResult r = run_path_edits(jobs, sorted, (uint32_t*)(uintptr_t)dag->nodes, (uint64_t*)(uintptr_t)dag->leaves,
                                 (const uint32_t*)(uintptr_t)dag->attrs, (uint32_t*)(uintptr_t)attrs_out,
                                 (hipStream_t)stream);
    if (r != PYR_OK) return r;
    *out_root = jobs[0].new_root;
    *out_count = jobs[0].new_count;
    *out_flags = jobs[0].status;
    return PYR_OK;
}

}  // extern "C"
