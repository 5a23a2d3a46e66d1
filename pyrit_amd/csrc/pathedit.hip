// pathedit.hip - voxel edits applied to a finished DAG in place of a rebuild
// (the reference's src/device/pathedit.zig, driven by context.zig:1093+ and
// world.zig:394-460).
//
// An edit touches only the path from the root to its brick: the brick and
// the nodes above it are written again as copies with the changed mask,
// voxel count and child reference; everything else stays shared. The copies
// go to an arena above every batch in the pools, so the chunk's ordinary
// relative references reach them and neither the format nor the traversal
// changes. Attributes are in DFS rank order: an inserted or removed voxel
// shifts every later rank by one.
//
// One block per job (chunk). Thread 0 walks the tree once per brick (the
// host groups a job's edits by brick, keeping their order) and turns every
// edit into (operation, rank, value) in the evolving rank space; then all
// threads of the block write the new attribute array: each output rank is
// traced back through the operations, newest first, to the edit that wrote
// it or to its rank in the old array. The walk is a chain of dependent
// loads (latency-bound by nature); many jobs run side by side.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

constexpr int kPathBlock = 256;
enum : uint32_t { kOpNone = 0, kOpSet = 1, kOpInsert = 2, kOpRemove = 3 };

__device__ inline uint32_t child_index(const uint32_t v[3], uint32_t s) {
    return ((v[0] >> s) & 1u) | (((v[1] >> s) & 1u) << 1) | (((v[2] >> s) & 1u) << 2);
}

__global__ void __launch_bounds__(kPathBlock) k_path_edit(PathEditJob* jobs, uint32_t njobs, uint4* edits,
                                                           uint32_t* nodes_pool, uint64_t* leaves_pool,
                                                           const uint32_t* src_pool, uint32_t* dst_pool) {
    __shared__ uint32_t s_count, s_flags;
    const uint32_t jid = blockIdx.x;
    if (jid >= njobs) return;
    PathEditJob& job = jobs[jid];
    uint32_t* nodes = nodes_pool + job.node_base;
    uint64_t* leaves = leaves_pool + job.leaf_base;
    uint4* ed = edits + job.edit_first;
    const uint32_t ne = job.edit_count;
    if (threadIdx.x == 0) {
        const uint32_t L = job.log2_size;
        uint32_t root = job.root, count = job.attr_count, flags = 0;
        uint32_t node_next = job.node_arena - job.node_base;
        const uint32_t node_end = node_next + job.node_cap;
        uint32_t leaf_next = job.leaf_arena - job.leaf_base;
        const uint32_t leaf_end = leaf_next + job.leaf_cap;
        uint32_t e = 0;
        while (e < ne) {
            const uint4 lead = ed[e];
            uint32_t end = e + 1;
            while (end < ne && (ed[end].x >> 2) == (lead.x >> 2) && (ed[end].y >> 2) == (lead.y >> 2) &&
                   (ed[end].z >> 2) == (lead.z >> 2))
                end++;
            if (flags) {  // the job already failed: the chunk is rebuilt instead
                for (; e < end; e++) ed[e] = make_uint4(kOpNone, 0, 0, 0);
                continue;
            }
            const uint32_t v[3] = {lead.x, lead.y, lead.z};
            // path to the brick of v: path[s] = node whose children have edge
            // 2^s; `missing` = level whose child on the path is absent (0: the
            // brick exists); prefix = voxels before the brick in DFS order
            uint32_t path[21];
            uint32_t missing = 0, prefix = 0;
            uint64_t brick = 0;
            {
                uint32_t node = root, s = L - 1;
                for (;;) {
                    path[s] = node;
                    const uint32_t mask = nodes[node] & 0xFFu;
                    const uint32_t ci = child_index(v, s);
                    const uint32_t before = (uint32_t)__popc(mask & ((1u << ci) - 1u));
                    for (uint32_t k = 0; k < before; k++) {
                        const uint32_t ref = nodes[node + 2 + k];
                        prefix += s == 2 ? (uint32_t)__popcll(leaves[ref]) : nodes[ref + 1];
                    }
                    if (!(mask & (1u << ci))) {
                        missing = s;
                        break;
                    }
                    const uint32_t ref = nodes[node + 2 + before];
                    if (s == 2) {
                        brick = leaves[ref];
                        break;
                    }
                    node = ref;
                    s--;
                }
            }
            // the brick's edits, in order: rank in the evolving brick
            uint64_t nb = brick;
            int32_t delta = 0;
            for (; e < end; e++) {
                const uint4 q = ed[e];
                const uint32_t bit = (q.x & 3u) + 4u * (q.y & 3u) + 16u * (q.z & 3u);
                const uint64_t bm = 1ull << bit;
                const uint32_t rank = prefix + (uint32_t)__popcll(nb & (bm - 1));
                const bool has = (nb & bm) != 0;
                uint32_t op = kOpNone;
                if (q.w != 0 && has) {
                    op = kOpSet;
                } else if (q.w != 0) {
                    op = kOpInsert;
                    nb |= bm;
                    delta++;
                } else if (has) {
                    op = kOpRemove;
                    nb &= ~bm;
                    delta--;
                }
                ed[e] = make_uint4(op, rank, q.w, 0);
            }
            if (nb == brick) continue;
            // new brick and new nodes up to the root
            int64_t child = -1;
            if (nb) {
                if (leaf_next >= leaf_end) {
                    flags |= kPathOverflow;
                    continue;
                }
                leaves[leaf_next] = nb;
                child = leaf_next++;
            }
            for (uint32_t s = 2; s < L && !(flags & kPathOverflow); s++) {
                const uint32_t ci = child_index(v, s);
                const bool exists = missing == 0 || s >= missing;
                uint32_t mask = 0, cnt = 0, n = 0, refs[8];
                if (exists) {
                    const uint32_t old = path[s];
                    const uint32_t om = nodes[old] & 0xFFu;
                    cnt = nodes[old + 1] + (uint32_t)delta;
                    uint32_t slot = 0;
                    for (uint32_t k = 0; k < 8; k++) {
                        const uint32_t kb = 1u << k;
                        if (k == ci) {
                            if (child >= 0) {
                                refs[n++] = (uint32_t)child;
                                mask |= kb;
                            }
                            if (om & kb) slot++;
                        } else if (om & kb) {
                            refs[n++] = nodes[old + 2 + slot];
                            slot++;
                            mask |= kb;
                        }
                    }
                } else if (child >= 0) {
                    mask = 1u << ci;
                    refs[n++] = (uint32_t)child;
                    cnt = (uint32_t)delta;
                }
                if (mask == 0) {
                    child = -1;
                    continue;
                }
                if (node_next + 2 + n > node_end) {
                    flags |= kPathOverflow;
                    break;
                }
                nodes[node_next] = mask;
                nodes[node_next + 1] = cnt;
                for (uint32_t k = 0; k < n; k++) nodes[node_next + 2 + k] = refs[k];
                child = node_next;
                node_next += 2 + n;
            }
            if (flags & kPathOverflow) continue;
            if (child < 0) flags |= kPathEmpty;
            else root = (uint32_t)child;
            count = (uint32_t)((int64_t)count + delta);
        }
        job.out_root = root;
        job.out_count = count;
        job.out_flags = flags;
        s_count = count;
        s_flags = flags;
    }
    __syncthreads();
    if (s_flags) return;  // failed jobs are rebuilt; their attributes are not used
    // attributes: output rank i traced back through the operations
    const uint32_t nf = s_count;
    const uint32_t* src = src_pool + job.attr_src;
    const uint32_t* pal = job.palette ? src_pool + job.pal_src : nullptr;
    uint32_t* dst = dst_pool + job.attr_dst;
    for (uint32_t i = threadIdx.x; i < nf; i += kPathBlock) {
        uint32_t idx = i, val = 0;
        bool found = false;
        for (uint32_t k = ne; k-- > 0;) {
            const uint4 o = ed[k];
            if (o.x == kOpInsert) {
                if (idx == o.y) {
                    val = o.z;
                    found = true;
                    break;
                }
                if (idx > o.y) idx--;
            } else if (o.x == kOpRemove) {
                if (idx >= o.y) idx++;
            } else if (o.x == kOpSet && idx == o.y) {
                val = o.z;
                found = true;
                break;
            }
        }
        if (!found) val = pal ? pal[(src[idx >> 3] >> ((idx & 7u) * 4u)) & 15u] : src[idx];
        dst[i] = val;
    }
}

}  // namespace

PyrResult run_path_edits(std::vector<PathEditJob>& jobs, const std::vector<uint32_t>& edits, uint32_t* nodes_pool,
                         uint64_t* leaves_pool, const uint32_t* src_pool, uint32_t* dst_pool, hipStream_t s) {
    if (jobs.empty()) return PYR_OK;
    PathEditJob* d_jobs = nullptr;
    uint4* d_edits = nullptr;
    const uint64_t jb = jobs.size() * sizeof(PathEditJob), eb = std::max<uint64_t>(edits.size() * 4, 16);
    PyrResult r = PYR_OK;
    if (pool_alloc((void**)&d_jobs, jb, s) != hipSuccess || pool_alloc((void**)&d_edits, eb, s) != hipSuccess) {
        (void)hipGetLastError();
        r = fail(PYR_ERROR_OUT_OF_MEMORY, "path edit jobs");
    }
    if (r == PYR_OK && (hipMemcpyAsync(d_jobs, jobs.data(), jb, hipMemcpyHostToDevice, s) != hipSuccess ||
                        (!edits.empty() && hipMemcpyAsync(d_edits, edits.data(), edits.size() * 4,
                                                          hipMemcpyHostToDevice, s) != hipSuccess)))
        r = fail(PYR_ERROR_CUDA, "path edit upload");
    if (r == PYR_OK) {
        hipLaunchKernelGGL(k_path_edit, dim3((uint32_t)jobs.size()), dim3(kPathBlock), 0, s, d_jobs,
                           (uint32_t)jobs.size(), d_edits, nodes_pool, leaves_pool, src_pool, dst_pool);
        if (hipMemcpyAsync(jobs.data(), d_jobs, jb, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            r = fail(PYR_ERROR_CUDA, "path edit");
    }
    if (r != PYR_OK) (void)hipGetLastError();
    pool_free(d_jobs, s);
    pool_free(d_edits, s);
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
    std::vector<uint32_t> ed(edits, edits + (size_t)count * 4);
    for (uint32_t i = 0; i < count; i++)
        if (ed[i * 4] >= n || ed[i * 4 + 1] >= n || ed[i * 4 + 2] >= n)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "edit %u outside the DAG", i);
    // group by brick, keeping the order inside a brick
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    auto brick_of = [&](uint32_t i) {
        return ((uint64_t)(ed[i * 4] >> 2) << 42) | ((uint64_t)(ed[i * 4 + 1] >> 2) << 21) | (ed[i * 4 + 2] >> 2);
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return brick_of(a) < brick_of(b); });
    std::vector<uint32_t> sorted((size_t)count * 4);
    for (uint32_t i = 0; i < count; i++)
        for (int c = 0; c < 4; c++) sorted[(size_t)i * 4 + c] = ed[(size_t)order[i] * 4 + c];
    std::vector<PathEditJob> jobs(1);
    PathEditJob& j = jobs[0];
    j = PathEditJob();
    j.root = dag->root;
    j.log2_size = dag->log2_size;
    j.attr_src = 0;
    j.attr_count = voxel_count;
    j.node_arena = node_words;
    j.node_cap = node_capacity - node_words;
    j.leaf_arena = leaf_count;
    j.leaf_cap = leaf_capacity - leaf_count;
    j.edit_first = 0;
    j.edit_count = count;
    j.attr_dst = 0;
    PyrResult r = run_path_edits(jobs, sorted, (uint32_t*)(uintptr_t)dag->nodes, (uint64_t*)(uintptr_t)dag->leaves,
                                 (const uint32_t*)(uintptr_t)dag->attrs, (uint32_t*)(uintptr_t)attrs_out,
                                 (hipStream_t)stream);
    if (r != PYR_OK) return r;
    *out_root = jobs[0].out_root;
    *out_count = jobs[0].out_count;
    *out_flags = jobs[0].out_flags;
    return PYR_OK;
}

}  // extern "C"
