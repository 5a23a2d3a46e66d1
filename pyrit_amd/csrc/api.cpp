// api.cpp - the pyrit.h context layer on the HIP runtime: geometry pools,
// GPU builds installed into them, instances, frame commit and the trace /
// render entry points (the hot-path subset of the reference's
// src/context.zig and src/root.zig exports).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "internal.h"
#include "range_alloc.h"

namespace pyr {

thread_local char g_diag[512];

// levels: the largest log2_size of a geometry the committed scene references
// (the per-ray LDS stack is sized to it); skipping: a material set may apply
// to these rays (the skip flag with transparent materials, or cutout
// materials), which selects the walk that checks them
PyrResult launch_trace_scene(const PyrScene* scene_dev, uint64_t rays, uint64_t hits, uint32_t count,
                             uint32_t ray_mask, uint32_t flags, uint32_t levels, bool skipping, hipStream_t s);
PyrResult launch_render_scene(const PyrScene* scene_dev, const PyrCamera& cam, const float* w2v_dev, uint64_t hits,
                              uint64_t depth, uint32_t ray_mask, uint32_t flags, uint32_t levels, bool skipping,
                              hipStream_t s);

namespace {

struct Geometry {
    bool live = false;
    uint32_t gen = 0;
    PyrGeometryData data{};
    uint64_t node_words = 0, leaf_count = 0, attr_count = 0;
    uint64_t* keys = nullptr;  // editable: sorted unique voxels (device)
    uint32_t* vals = nullptr;
    uint32_t voxel_count = 0;
    uint32_t users = 0;  // instances referring to it
    int32_t shared = -1; // >= 0: pool ranges of a chunk batch shared with other geometries
    // after a path edit: nodes also in the path-edit arena, attributes in a range of its own
    bool uses_arena = false;
    uint64_t own_attr_offset = 0, own_attr_count = 0;
    // destroyed, its ranges not yet released: the slot is not reused before
    // the release, so the committed geometry table keeps its entry valid
    bool held = false;
};

// Pool ranges and buffers that the committed scene may still reach: the
// previous version of an edited geometry, a destroyed geometry, attributes
// a path edit replaced. Traces and renders launched before the next
// pyr_commit still read them (docs/API.md:91), so they are released by that
// commit, once its table uploads and everything queued before them on the
// context stream have finished (context.zig:972-1005, 1446-1545).
struct Deferred {
    uint64_t node_offset = 0, node_words = 0;
    uint64_t leaf_offset = 0, leaf_count = 0;
    uint64_t attr_offset = 0, attr_count = 0;
    uint64_t own_attr_offset = 0, own_attr_count = 0;
    int32_t shared = -1;      // a batch reference to drop
    bool arena_user = false;  // a path-edit arena user to drop
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    int32_t slot = -1;        // a destroyed geometry's slot, free again after the release
};

// pool ranges of one installed chunk batch; every chunk's geometry refers to
// the batch's node and leaf ranges and to its own slice of the attributes
struct SharedRange {
    uint64_t node_offset = 0, leaf_offset = 0, attr_offset = 0;
    uint64_t node_words = 0, leaf_count = 0, attr_count = 0;
    uint32_t refs = 0;
};

struct Instance {
    bool live = false;
    uint32_t gen = 0;
    PyrInstanceData data{};
};

struct Pool {
    void* base = nullptr;
    uint64_t elem = 0;  // bytes per element
    RangeAlloc ra{0};
    // path-edit arena: the top eighth of the node and leaf pools, above every
    // range the allocator hands out, bump-allocated (context.zig:435-443)
    uint64_t arena_start = 0, arena_next = 0, arena_end = 0;
};

}  // namespace

// 3x4 affine inverse in f64 (the reference's xform.zig:8-34)
bool inverse34(const float m[12], float out[12]) {
    const double a[9] = {m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]};
    const double t[3] = {m[3], m[7], m[11]};
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = a[5] * a[6] - a[3] * a[8];
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(fabs(det) > 1e-30) || !isfinite(det)) return false;
    const double id = 1.0 / det;
    const double r[9] = {
        c00 * id, (a[2] * a[7] - a[1] * a[8]) * id, (a[1] * a[5] - a[2] * a[4]) * id,
        c01 * id, (a[0] * a[8] - a[2] * a[6]) * id, (a[2] * a[3] - a[0] * a[5]) * id,
        c02 * id, (a[1] * a[6] - a[0] * a[7]) * id, (a[0] * a[4] - a[1] * a[3]) * id,
    };
    for (int row = 0; row < 3; row++) {
        const double r0 = r[row * 3], r1 = r[row * 3 + 1], r2 = r[row * 3 + 2];
        out[row * 4 + 0] = (float)r0;
        out[row * 4 + 1] = (float)r1;
        out[row * 4 + 2] = (float)r2;
        out[row * 4 + 3] = (float)(-(r0 * t[0] + r1 * t[1] + r2 * t[2]));
    }
    return true;
}

// world AABB of [0, size]^3 under m with a 1e-5 relative margin (xform.zig:37-58)
void transformed_box(const float m[12], float size, float lo[3], float hi[3]) {
    for (int row = 0; row < 3; row++) {
        lo[row] = m[row * 4 + 3];
        hi[row] = m[row * 4 + 3];
        for (int col = 0; col < 3; col++) {
            const float e = m[row * 4 + col] * size;
            lo[row] += fminf(e, 0.0f);
            hi[row] += fmaxf(e, 0.0f);
        }
    }
    for (int i = 0; i < 3; i++) {
        const float eps = 1e-5f * fmaxf(1.0f, fmaxf(fabsf(lo[i]), fabsf(hi[i])));
        lo[i] -= eps;
        hi[i] += eps;
    }
}

}  // namespace pyr

using namespace pyr;

struct PyrContext {
    int device = 0;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    Pool nodes, leaves, attrs;
    std::vector<Geometry> geos;
    std::vector<SharedRange> shared;
    std::vector<Instance> insts;
    uint32_t max_instances = 0, max_geometries = 0;
    uint64_t frame = 0;
    double time = 0, time_prev = 0;
    double origin[3] = {0, 0, 0}, origin_prev[3] = {0, 0, 0};
    // committed device state
    PyrGeometryData* geo_dev = nullptr;
    PyrInstanceData* inst_dev[2] = {nullptr, nullptr};  // current, previous
    PyrScene* scene_dev = nullptr;
    float* w2v_dev = nullptr;
    uint32_t history = 0;
    uint32_t views = 0;
    uint32_t scene_levels = 3;  // largest geometry edge (log2) referenced by the committed instances
    // material table (context.zig:1948-1964): host copy for the transparent /
    // cutout sets of the next commit, device copy for the caller's kernels
    std::vector<PyrMaterial> materials;
    PyrMaterial* materials_dev = nullptr;
    bool scene_transparent = false, scene_cutout = false;  // non-empty sets in the committed scene
    uint32_t arena_users = 0;  // geometries with nodes in the path-edit arena (deferred ones included)
    std::vector<Deferred> deferred;
};

namespace {

// handles: bits 0..31 index + 1, bits 32..63 generation (context.zig:36-43)
inline uint64_t make_handle(uint32_t index, uint32_t gen) { return ((uint64_t)gen << 32) | (uint64_t)(index + 1); }
inline uint32_t handle_index(uint64_t h) { return (uint32_t)(h & 0xFFFFFFFFu) - 1; }
inline uint32_t handle_gen(uint64_t h) { return (uint32_t)(h >> 32); }

Geometry* geo_of(PyrContext* ctx, PyrGeometry g) {
    uint64_t h = (uint64_t)(uintptr_t)g;
    uint32_t i = handle_index(h);
    if (!h || i >= ctx->geos.size()) return nullptr;
    Geometry& ge = ctx->geos[i];
    return ge.live && ge.gen == handle_gen(h) ? &ge : nullptr;
}

Instance* inst_of(PyrContext* ctx, PyrInstance in) {
    uint64_t h = (uint64_t)(uintptr_t)in;
    uint32_t i = handle_index(h);
    if (!h || i >= ctx->insts.size()) return nullptr;
    Instance& it = ctx->insts[i];
    return it.live && it.gen == handle_gen(h) ? &it : nullptr;
}

PyrResult hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return PYR_OK;
    (void)hipGetLastError();
    return fail(PYR_ERROR_CUDA, "%s: %s", what, hipGetErrorString(e));
}

PyrResult pool_init(Pool& p, uint64_t bytes, uint64_t elem, bool arena) {
    p.elem = elem;
    const uint64_t total = bytes / elem;
    p.arena_end = total;
    p.arena_start = p.arena_next = arena ? total - total / 8 : total;
    p.ra = RangeAlloc(p.arena_start);
    return hip_check(hipMalloc(&p.base, bytes), "pool allocation");
}

void release_shared(PyrContext* ctx, int32_t id) {
    SharedRange& r = ctx->shared[(size_t)id];
    if (r.refs == 0 || --r.refs > 0) return;
    ctx->nodes.ra.release(r.node_offset, r.node_words);
    ctx->leaves.ra.release(r.leaf_offset, r.leaf_count);
    ctx->attrs.ra.release(r.attr_offset, r.attr_count);
    r = SharedRange();
}

// everything a geometry holds in the pools and on the device, moved into a
// deferred release; g keeps its table entry (data) but owns nothing
Deferred take_geometry(Geometry& g) {
    Deferred d;
    d.own_attr_offset = g.own_attr_offset;
    d.own_attr_count = g.own_attr_count;
    d.arena_user = g.uses_arena;
    d.shared = g.shared;
    if (g.shared < 0) {
        d.node_offset = g.data.node_offset;
        d.node_words = g.node_words;
        d.leaf_offset = g.data.leaf_offset;
        d.leaf_count = g.leaf_count;
        d.attr_offset = g.data.attribute_offset;
        d.attr_count = g.attr_count;
    }
    d.keys = g.keys;
    d.vals = g.vals;
    g.own_attr_offset = g.own_attr_count = 0;
    g.uses_arena = false;
    g.shared = -1;
    g.node_words = g.leaf_count = g.attr_count = 0;
    g.keys = nullptr;
    g.vals = nullptr;
    return d;
}

void release_deferred(PyrContext* ctx, const Deferred& d) {
    if (d.node_words) ctx->nodes.ra.release(d.node_offset, d.node_words);
    if (d.leaf_count) ctx->leaves.ra.release(d.leaf_offset, d.leaf_count);
    if (d.attr_count) ctx->attrs.ra.release(d.attr_offset, d.attr_count);
    if (d.own_attr_count) ctx->attrs.ra.release(d.own_attr_offset, d.own_attr_count);
    if (d.shared >= 0) release_shared(ctx, d.shared);
    if (d.arena_user) ctx->arena_users--;
    pool_free(d.keys, ctx->stream);
    pool_free(d.vals, ctx->stream);
    if (d.slot >= 0) ctx->geos[(size_t)d.slot].held = false;
}

// a geometry slot that is neither live nor held by a deferred release
uint32_t free_geometry_slot(PyrContext* ctx) {
    uint32_t slot = 0;
    while (slot < ctx->geos.size() && (ctx->geos[slot].live || ctx->geos[slot].held)) slot++;
    return slot;
}

// copy a finished build into the pools (context.zig:930-952, installBuilt)
PyrResult install(PyrContext* ctx, Geometry& g, const PyrhipBuilt& b, bool editable) {
    int64_t no = ctx->nodes.ra.alloc(b.node_words);
    int64_t lo = ctx->leaves.ra.alloc(b.leaf_count);
    int64_t ao = ctx->attrs.ra.alloc(b.voxel_count);
    if (no < 0 || lo < 0 || ao < 0) {
        if (no >= 0) ctx->nodes.ra.release((uint64_t)no, b.node_words);
        if (lo >= 0) ctx->leaves.ra.release((uint64_t)lo, b.leaf_count);
        if (ao >= 0) ctx->attrs.ra.release((uint64_t)ao, b.voxel_count);
        return fail(PYR_ERROR_OUT_OF_MEMORY, "geometry pool full (%u words, %u leaves, %u attributes)", b.node_words,
                    b.leaf_count, b.voxel_count);
    }
    hipStream_t s = ctx->stream;
    PyrResult r = hip_check(hipMemcpyAsync((char*)ctx->nodes.base + (uint64_t)no * 4, (void*)(uintptr_t)b.nodes,
                                           (uint64_t)b.node_words * 4, hipMemcpyDeviceToDevice, s),
                            "node copy");
    if (r == PYR_OK)
        r = hip_check(hipMemcpyAsync((char*)ctx->leaves.base + (uint64_t)lo * 8, (void*)(uintptr_t)b.leaves,
                                     (uint64_t)b.leaf_count * 8, hipMemcpyDeviceToDevice, s),
                      "leaf copy");
    if (r == PYR_OK)
        r = hip_check(hipMemcpyAsync((char*)ctx->attrs.base + (uint64_t)ao * 4, (void*)(uintptr_t)b.attrs,
                                     (uint64_t)b.voxel_count * 4, hipMemcpyDeviceToDevice, s),
                      "attribute copy");
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(s), "install");
    if (r != PYR_OK) {
        ctx->nodes.ra.release((uint64_t)no, b.node_words);
        ctx->leaves.ra.release((uint64_t)lo, b.leaf_count);
        ctx->attrs.ra.release((uint64_t)ao, b.voxel_count);
        return r;
    }
    g.data.node_offset = (uint32_t)no;
    g.data.leaf_offset = (uint32_t)lo;
    g.data.attribute_offset = (uint32_t)ao;
    g.data.root = b.root;
    g.data.log2_size = b.log2_size;
    g.data.flags = PYR_GEOMETRY_HAS_ATTRIBUTES;
    g.data.default_attribute = 1;
    g.data.palette_offset = 0;
    g.node_words = b.node_words;
    g.leaf_count = b.leaf_count;
    g.attr_count = b.voxel_count;
    g.voxel_count = b.voxel_count;
    if (editable) {
        g.keys = (uint64_t*)(uintptr_t)b.keys;
        g.vals = (uint32_t*)(uintptr_t)b.attrs;
    } else {
        pool_free((void*)(uintptr_t)b.keys, s);
        pool_free((void*)(uintptr_t)b.attrs, s);
    }
    pool_free((void*)(uintptr_t)b.nodes, s);
    pool_free((void*)(uintptr_t)b.leaves, s);
    return PYR_OK;
}

// device copy of host voxels when PYR_BUILD_HOST_INPUT is set
PyrResult stage_voxels(PyrContext* ctx, const void* voxels, uint32_t count, uint32_t flags, void** dev, bool* owned) {
    *owned = false;
    *dev = (void*)voxels;
    if (!(flags & PYR_BUILD_HOST_INPUT) || count == 0) return PYR_OK;
    PyrResult r = hip_check(pool_alloc(dev, (uint64_t)count * sizeof(PyrVoxel), ctx->stream), "voxel upload buffer");
    if (r != PYR_OK) return r;
    *owned = true;
    return hip_check(hipMemcpyAsync(*dev, voxels, (uint64_t)count * sizeof(PyrVoxel), hipMemcpyHostToDevice,
                                    ctx->stream),
                     "voxel upload");
}

}  // namespace

// ---- chunk batches for streamed worlds (context.zig:1287+ installChunkBatch) ----

namespace pyr {

hipStream_t ctx_stream(PyrContext* ctx) { return ctx->stream; }

// Copies a finished chunk batch into the pools once (nodes, leaves and the
// attributes of every chunk, palette packed where a chunk has at most 16
// distinct values, context.zig:1311-1370; PYRIT_NO_PALETTE in the environment
// keeps 32 bits everywhere, as in the reference); on success the batch
// buffers are released and *id names the shared ranges, held by one
// reference until ctx_release_batch.
PyrResult ctx_install_batch(PyrContext* ctx, PyrhipBuilt* b, const std::vector<uint32_t>& first, int32_t* id,
                            std::vector<uint64_t>* attr_word, std::vector<uint8_t>* palette) {
    hipStream_t s = ctx->stream;
    const uint32_t k = (uint32_t)first.size() - 1;
    const uint32_t* attrs = (const uint32_t*)(uintptr_t)b->attrs;
    std::vector<uint32_t> counts(k, kPaletteOverflow);
    PaletteScratch ps;
    PyrResult r = PYR_OK;
    if (!getenv("PYRIT_NO_PALETTE")) r = palette_scan(attrs, first, s, &counts, &ps);
    if (r != PYR_OK) {
        palette_release(&ps, s);
        return r;
    }
    attr_word->assign(k, 0);
    palette->assign(k, 0);
    uint64_t words = 0;
    for (uint32_t j = 0; j < k; j++) {
        (*attr_word)[j] = words;
        (*palette)[j] = counts[j] != kPaletteOverflow;
        words += palette_words(counts[j], first[j + 1] - first[j]);
    }
    const int64_t no = ctx->nodes.ra.alloc(b->node_words);
    const int64_t lo = ctx->leaves.ra.alloc(b->leaf_count);
    const int64_t ao = ctx->attrs.ra.alloc(words);
    auto release = [&]() {
        if (no >= 0) ctx->nodes.ra.release((uint64_t)no, b->node_words);
        if (lo >= 0) ctx->leaves.ra.release((uint64_t)lo, b->leaf_count);
        if (ao >= 0) ctx->attrs.ra.release((uint64_t)ao, words);
    };
    if (no < 0 || lo < 0 || ao < 0) {
        release();
        palette_release(&ps, s);
        return fail(PYR_ERROR_OUT_OF_MEMORY, "geometry pool full (%u words, %u leaves, %llu attribute words)",
                    b->node_words, b->leaf_count, (unsigned long long)words);
    }
    r = hip_check(hipMemcpyAsync((char*)ctx->nodes.base + (uint64_t)no * 4, (void*)(uintptr_t)b->nodes,
                                 (uint64_t)b->node_words * 4, hipMemcpyDeviceToDevice, s),
                  "node copy");
    if (r == PYR_OK)
        r = hip_check(hipMemcpyAsync((char*)ctx->leaves.base + (uint64_t)lo * 8, (void*)(uintptr_t)b->leaves,
                                     (uint64_t)b->leaf_count * 8, hipMemcpyDeviceToDevice, s),
                      "leaf copy");
    if (r == PYR_OK) {
        if (ps.lo) {
            std::vector<uint64_t> dst(k);
            for (uint32_t j = 0; j < k; j++) dst[j] = (uint64_t)ao + (*attr_word)[j];
            r = palette_pack(attrs, first, dst, (uint32_t*)ctx->attrs.base, s, &ps);
        } else if (words) {
            r = hip_check(hipMemcpyAsync((char*)ctx->attrs.base + (uint64_t)ao * 4, attrs, words * 4,
                                         hipMemcpyDeviceToDevice, s),
                          "attribute copy");
        }
    }
    palette_release(&ps, s);
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(s), "install");
    if (r != PYR_OK) {
        release();
        return r;
    }
    size_t slot = 0;
    while (slot < ctx->shared.size() && ctx->shared[slot].refs) slot++;
    if (slot == ctx->shared.size()) ctx->shared.emplace_back();
    SharedRange& sr = ctx->shared[slot];
    sr.node_offset = (uint64_t)no;
    sr.leaf_offset = (uint64_t)lo;
    sr.attr_offset = (uint64_t)ao;
    sr.node_words = b->node_words;
    sr.leaf_count = b->leaf_count;
    sr.attr_count = words;
    sr.refs = 1;
    pyrhip_built_free(b, s);
    *id = (int32_t)slot;
    return PYR_OK;
}

void ctx_release_batch(PyrContext* ctx, int32_t id) {
    if (id >= 0 && (size_t)id < ctx->shared.size()) release_shared(ctx, id);
}

uint64_t ctx_batch_bytes(PyrContext* ctx, int32_t id) {
    const SharedRange& r = ctx->shared[(size_t)id];
    return r.node_words * 4 + r.leaf_count * 8 + r.attr_count * 4;
}

// A geometry for one chunk of an installed batch: the batch's node and leaf
// ranges (word refs are relative to the batch start), the chunk's root and
// its attributes at word attr_word of the batch's attribute range: 16
// palette entries and the index words (palette), or 32-bit values.
PyrResult ctx_geometry_from_batch(PyrContext* ctx, int32_t id, uint32_t root, uint64_t attr_word, bool palette,
                                  uint32_t voxel_count, uint32_t log2_size, PyrGeometry* out) {
    uint32_t slot = free_geometry_slot(ctx);
    if (slot >= ctx->max_geometries) return fail(PYR_ERROR_CAPACITY, "max_geometries reached");
    if (slot == ctx->geos.size()) ctx->geos.emplace_back();
    SharedRange& sr = ctx->shared[(size_t)id];
    Geometry& g = ctx->geos[slot];
    const uint32_t gen = g.gen + 1;
    g = Geometry();
    g.gen = gen;
    g.data.node_offset = (uint32_t)sr.node_offset;
    g.data.leaf_offset = (uint32_t)sr.leaf_offset;
    g.data.attribute_offset = (uint32_t)(sr.attr_offset + attr_word);
    g.data.flags = PYR_GEOMETRY_HAS_ATTRIBUTES;
    if (palette) {  // context.zig:1385-1389
        g.data.flags |= PYR_GEOMETRY_PALETTE;
        g.data.palette_offset = g.data.attribute_offset;
        g.data.attribute_offset += 16;
    }
    g.data.root = root;
    g.data.log2_size = log2_size;
    g.data.default_attribute = 1;
    g.voxel_count = voxel_count;
    g.shared = id;
    sr.refs++;
    g.live = true;
    *out = (PyrGeometry)(uintptr_t)make_handle(slot, g.gen);
    return PYR_OK;
}

PyrResult ctx_path_edit(PyrContext* ctx, const std::vector<PathEditRequest>& reqs, const std::vector<uint32_t>& edits,
                        std::vector<uint8_t>* applied, std::vector<uint32_t>* voxels) {
    applied->assign(reqs.size(), 0);
    voxels->assign(reqs.size(), 0);
    hipStream_t s = ctx->stream;
    // nothing refers to the arena: start it over (after the work in flight)
    if (ctx->arena_users == 0 &&
        (ctx->nodes.arena_next != ctx->nodes.arena_start || ctx->leaves.arena_next != ctx->leaves.arena_start)) {
        PyrResult r = hip_check(hipStreamSynchronize(s), "path edit");
        if (r != PYR_OK) return r;
        ctx->nodes.arena_next = ctx->nodes.arena_start;
        ctx->leaves.arena_next = ctx->leaves.arena_start;
    }
    std::vector<PathEditJob> jobs;
    std::vector<size_t> job_req;
    std::vector<uint64_t> caps;
    std::vector<uint32_t> sorted;
    for (size_t i = 0; i < reqs.size(); i++) {
        const PathEditRequest& q = reqs[i];
        Geometry* g = geo_of(ctx, q.geometry);
        // only chunks of installed batches: their node and leaf ranges lie below the arena
        if (!g || g->shared < 0 || q.count == 0 || q.count > kPathEditMax) continue;
        if (!(g->data.flags & PYR_GEOMETRY_HAS_ATTRIBUTES) || g->data.log2_size < 3) continue;
        if ((uint64_t)q.first + q.count > edits.size() / 4) continue;
        const uint32_t* e = &edits[(size_t)q.first * 4];
        const uint32_t n = 1u << g->data.log2_size;
        bool in_range = true;
        for (uint32_t k = 0; in_range && k < q.count; k++) in_range = e[k * 4] < n && e[k * 4 + 1] < n && e[k * 4 + 2] < n;
        if (!in_range) continue;
        const size_t at = sorted.size();
        sort_path_edits(e, q.count, &sorted);
        uint64_t words = 0, leaves = 0;
        path_edit_capacity(&sorted[at], q.count, g->data.log2_size, &words, &leaves);
        const uint64_t most = (uint64_t)g->voxel_count + q.count;  // attributes after the edits, at most
        const int64_t a = ctx->nodes.arena_next + words <= ctx->nodes.arena_end &&
                                  ctx->leaves.arena_next + leaves <= ctx->leaves.arena_end
                              ? ctx->attrs.ra.alloc(most)
                              : -1;
        if (a < 0) {
            sorted.resize(at);
            continue;
        }
        PathEditJob j;
        j.root = g->data.root;
        j.log2_size = g->data.log2_size;
        j.node_base = g->data.node_offset;
        j.leaf_base = g->data.leaf_offset;
        j.node_arena = (uint32_t)ctx->nodes.arena_next;
        j.node_end = (uint32_t)(ctx->nodes.arena_next + words);
        j.leaf_arena = (uint32_t)ctx->leaves.arena_next;
        j.leaf_end = (uint32_t)(ctx->leaves.arena_next + leaves);
        j.attr_src = g->data.attribute_offset;
        j.attr_count = g->voxel_count;
        if (g->data.flags & PYR_GEOMETRY_PALETTE) j.palette_src = g->data.palette_offset;
        j.attr_dst = (uint32_t)a;
        j.edit_first = (uint32_t)(at / 4);
        j.edit_count = q.count;
        ctx->nodes.arena_next += words;
        ctx->leaves.arena_next += leaves;
        jobs.push_back(j);
        job_req.push_back(i);
        caps.push_back(most);
    }
    if (jobs.empty()) return PYR_OK;
    uint32_t* attrs = (uint32_t*)ctx->attrs.base;
    PyrResult r = run_path_edits(jobs, sorted, (uint32_t*)ctx->nodes.base, (uint64_t*)ctx->leaves.base, attrs, attrs, s);
    for (size_t k = 0; k < jobs.size(); k++) {
        const PathEditJob& j = jobs[k];
        Geometry* g = geo_of(ctx, reqs[job_req[k]].geometry);
        if (r != PYR_OK || j.status || !g) {
            ctx->attrs.ra.release(j.attr_dst, caps[k]);
            continue;
        }
        if (g->own_attr_count) {  // the committed scene may still read them
            Deferred d;
            d.own_attr_offset = g->own_attr_offset;
            d.own_attr_count = g->own_attr_count;
            ctx->deferred.push_back(d);
        }
        g->own_attr_offset = j.attr_dst;
        g->own_attr_count = caps[k];
        g->data.root = j.new_root;
        g->data.attribute_offset = j.attr_dst;
        g->data.flags = PYR_GEOMETRY_HAS_ATTRIBUTES;  // 32-bit attributes ...
        g->data.palette_offset = 0;
        g->voxel_count = j.new_count;
        // ... unless the chunk was palette packed and its edited attributes
        // still take at most 16 values: packed again (palette.zig:1-152)
        if (j.palette_src != 0xFFFFFFFFu && j.new_count && !getenv("PYRIT_NO_PALETTE")) {
            const std::vector<uint32_t> lo = {0u, j.new_count};
            std::vector<uint32_t> counts;
            PaletteScratch ps;
            PyrResult pr = palette_scan(attrs + j.attr_dst, lo, s, &counts, &ps);
            if (pr == PYR_OK && counts[0] != kPaletteOverflow) {
                const uint64_t words = palette_words(counts[0], j.new_count);
                const int64_t pa = ctx->attrs.ra.alloc(words);
                if (pa >= 0 && palette_pack(attrs + j.attr_dst, lo, {(uint64_t)pa}, attrs, s, &ps) == PYR_OK) {
                    // (the 32-bit range is reused only by later work on this stream)
                    ctx->attrs.ra.release(j.attr_dst, caps[k]);
                    g->own_attr_offset = (uint64_t)pa;
                    g->own_attr_count = words;
                    g->data.flags = PYR_GEOMETRY_HAS_ATTRIBUTES | PYR_GEOMETRY_PALETTE;
                    g->data.palette_offset = (uint32_t)pa;
                    g->data.attribute_offset = (uint32_t)pa + 16;
                } else if (pa >= 0) {
                    ctx->attrs.ra.release((uint64_t)pa, words);
                }
            }
            palette_release(&ps, s);
            (void)hipGetLastError();
            diag_clear();
        }
        if (!g->uses_arena) {
            g->uses_arena = true;
            ctx->arena_users++;
        }
        (*applied)[job_req[k]] = 1;
        (*voxels)[job_req[k]] = j.new_count;
    }
    return r;
}

bool ctx_arena_pressure(PyrContext* ctx) {
    const Pool& n = ctx->nodes;
    const Pool& l = ctx->leaves;
    return (n.arena_next - n.arena_start) * 4 > (n.arena_end - n.arena_start) * 3 ||
           (l.arena_next - l.arena_start) * 4 > (l.arena_end - l.arena_start) * 3;
}

}  // namespace pyr

extern "C" {

PYR_API const char* pyr_result_string(PyrResult r) {
    switch (r) {
        case PYR_OK: return "PYR_OK";
        case PYR_ERROR_INVALID_ARGUMENT: return "PYR_ERROR_INVALID_ARGUMENT";
        case PYR_ERROR_INVALID_HANDLE: return "PYR_ERROR_INVALID_HANDLE";
        case PYR_ERROR_OUT_OF_MEMORY: return "PYR_ERROR_OUT_OF_MEMORY";
        case PYR_ERROR_CAPACITY: return "PYR_ERROR_CAPACITY";
        case PYR_ERROR_IN_USE: return "PYR_ERROR_IN_USE";
        case PYR_ERROR_CUDA: return "PYR_ERROR_CUDA";
        case PYR_ERROR_COMPILE: return "PYR_ERROR_COMPILE";
        case PYR_ERROR_NOT_FOUND: return "PYR_ERROR_NOT_FOUND";
        case PYR_ERROR_VERSION: return "PYR_ERROR_VERSION";
        default: return "PYR_ERROR_UNKNOWN";
    }
}

PYR_API const char* pyr_error_message(void) { return g_diag; }

PYR_API PyrResult pyr_create(const PyrCreateInfo* info, PyrContext** out) {
    diag_clear();
    if (!info || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (info->struct_size != sizeof(PyrCreateInfo)) return fail(PYR_ERROR_VERSION, "PyrCreateInfo size mismatch");
    if ((info->version >> 16) != PYR_VERSION_MAJOR) return fail(PYR_ERROR_VERSION, "incompatible PYR_VERSION");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_NOT_FOUND, "no HIP device");
    }
    if (info->device < 0 || info->device >= n) return fail(PYR_ERROR_INVALID_ARGUMENT, "device %d", info->device);
    PyrContext* ctx = new PyrContext();
    ctx->device = info->device;
    ctx->flags = info->flags;
    PyrResult r = hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (r == PYR_OK) {
        if (info->cuda_stream) {
            ctx->stream = (hipStream_t)info->cuda_stream;
        } else {
            r = hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream");
            ctx->own_stream = r == PYR_OK;
        }
    }
    ctx->max_instances = info->max_instances ? info->max_instances : 65536;
    ctx->max_geometries = info->max_geometries ? info->max_geometries : 4096;
    const uint64_t mib = 1ull << 20;
    if (r == PYR_OK) r = pool_init(ctx->nodes, info->node_pool_bytes ? info->node_pool_bytes : 256 * mib, 4, true);
    if (r == PYR_OK) r = pool_init(ctx->leaves, info->leaf_pool_bytes ? info->leaf_pool_bytes : 256 * mib, 8, true);
    if (r == PYR_OK)
        r = pool_init(ctx->attrs, info->attribute_pool_bytes ? info->attribute_pool_bytes : 256 * mib, 4, false);
    if (r == PYR_OK)
        r = hip_check(hipMalloc(&ctx->geo_dev, (uint64_t)ctx->max_geometries * sizeof(PyrGeometryData)), "tables");
    for (int q = 0; q < 2 && r == PYR_OK; q++)
        r = hip_check(hipMalloc(&ctx->inst_dev[q], (uint64_t)ctx->max_instances * sizeof(PyrInstanceData)), "tables");
    if (r == PYR_OK) r = hip_check(hipMalloc(&ctx->scene_dev, sizeof(PyrScene)), "scene");
    if (r == PYR_OK) r = hip_check(hipMalloc(&ctx->w2v_dev, 64), "camera");
    if (r == PYR_OK) {
        PyrMaterial m;
        pyr_material_default(&m);
        ctx->materials.assign(PYR_MAX_MATERIALS, m);
        r = hip_check(hipMalloc(&ctx->materials_dev, PYR_MAX_MATERIALS * sizeof(PyrMaterial)), "materials");
    }
    if (r == PYR_OK)
        r = hip_check(hipMemcpyAsync(ctx->materials_dev, ctx->materials.data(), PYR_MAX_MATERIALS * sizeof(PyrMaterial),
                                     hipMemcpyHostToDevice, ctx->stream),
                      "materials");
    if (r == PYR_OK) r = hip_check(hipMemsetAsync(ctx->scene_dev, 0, sizeof(PyrScene), ctx->stream), "scene");
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(ctx->stream), "create");
    if (r != PYR_OK) {
        char msg[512];
        memcpy(msg, g_diag, sizeof(msg));
        pyr_destroy(ctx);
        memcpy(g_diag, msg, sizeof(msg));
        return r;
    }
    *out = ctx;
    return pyr_commit(ctx, nullptr);
}

PYR_API void pyr_destroy(PyrContext* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (Geometry& g : ctx->geos) {
        if (!g.live) continue;
        pool_free(g.keys, ctx->stream);
        pool_free(g.vals, ctx->stream);
    }
    for (const Deferred& d : ctx->deferred) {
        pool_free(d.keys, ctx->stream);
        pool_free(d.vals, ctx->stream);
    }
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->nodes.base);
    (void)hipFree(ctx->leaves.base);
    (void)hipFree(ctx->attrs.base);
    (void)hipFree(ctx->geo_dev);
    (void)hipFree(ctx->inst_dev[0]);
    (void)hipFree(ctx->inst_dev[1]);
    (void)hipFree(ctx->scene_dev);
    (void)hipFree(ctx->w2v_dev);
    (void)hipFree(ctx->materials_dev);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

PYR_API void* pyr_cuda_stream(PyrContext* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

PYR_API uint32_t pyr_features(PyrContext* ctx) {
    (void)ctx;
    return 0;
}

// stats (context.zig:2999-3017)
PYR_API PyrResult pyr_get_stats(PyrContext* ctx, PyrStats* out) {
    diag_clear();
    if (!ctx || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    PyrStats st{};
    st.node_pool_used = ctx->nodes.ra.used() * ctx->nodes.elem;
    st.node_pool_capacity = ctx->nodes.ra.capacity() * ctx->nodes.elem;
    st.leaf_pool_used = ctx->leaves.ra.used() * ctx->leaves.elem;
    st.leaf_pool_capacity = ctx->leaves.ra.capacity() * ctx->leaves.elem;
    st.attribute_pool_used = ctx->attrs.ra.used() * ctx->attrs.elem;
    st.attribute_pool_capacity = ctx->attrs.ra.capacity() * ctx->attrs.elem;
    for (const Geometry& g : ctx->geos) st.geometries += g.live ? 1u : 0u;
    for (const Instance& in : ctx->insts) st.instances += in.live ? 1u : 0u;
    st.frame = ctx->frame;
    st.features = 0;
    *out = st;
    return PYR_OK;
}

PYR_API PyrResult pyr_synchronize(PyrContext* ctx) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    return hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
}

// geometryBuild (context.zig:954-969)
PYR_API PyrResult pyr_geometry_build(PyrContext* ctx, uint32_t log2_size, const void* voxels, uint32_t count,
                                     uint32_t flags, PyrGeometry* out) {
    diag_clear();
    if (!ctx || !out || (count && !voxels)) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    uint32_t slot = free_geometry_slot(ctx);
    if (slot >= ctx->max_geometries) return fail(PYR_ERROR_CAPACITY, "max_geometries reached");
    void* dev = nullptr;
    bool owned = false;
    PyrResult r = stage_voxels(ctx, voxels, count, flags, &dev, &owned);
    PyrhipBuilt b{};
    if (r == PYR_OK) r = pyrhip_build(log2_size, 0, 0, 0, (uint64_t)(uintptr_t)dev, count, nullptr, ctx->stream, &b);
    if (owned) pool_free(dev, ctx->stream);
    if (r != PYR_OK) return r;
    if (slot == ctx->geos.size()) ctx->geos.emplace_back();
    Geometry& g = ctx->geos[slot];
    uint32_t gen = g.gen + 1;
    g = Geometry();
    g.gen = gen;
    r = install(ctx, g, b, (flags & PYR_BUILD_EDITABLE) != 0);
    if (r != PYR_OK) {
        pyrhip_built_free(&b, ctx->stream);
        return r;
    }
    g.live = true;
    *out = (PyrGeometry)(uintptr_t)make_handle(slot, g.gen);
    return PYR_OK;
}

// geometryEdit (context.zig:973-1006): rebuild from the kept sorted voxels
// plus the edits; the new version replaces the old in place
PYR_API PyrResult pyr_geometry_edit(PyrContext* ctx, PyrGeometry geometry, const void* voxels, uint32_t count,
                                    uint32_t flags) {
    diag_clear();
    if (!ctx || (count && !voxels)) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    Geometry* g = geo_of(ctx, geometry);
    if (!g) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    if (!g->keys) return fail(PYR_ERROR_INVALID_ARGUMENT, "geometry was not built with PYR_BUILD_EDITABLE");
    void* dev = nullptr;
    bool owned = false;
    PyrResult r = stage_voxels(ctx, voxels, count, flags, &dev, &owned);
    PyrhipBuilt b{};
    if (r == PYR_OK)
        r = pyrhip_build(g->data.log2_size, (uint64_t)(uintptr_t)g->keys, (uint64_t)(uintptr_t)g->vals,
                         g->voxel_count, (uint64_t)(uintptr_t)dev, count, nullptr, ctx->stream, &b);
    if (owned) pool_free(dev, ctx->stream);
    if (r != PYR_OK) return r;
    // the new version goes into fresh ranges; the old one stays readable by
    // the committed scene until the next pyr_commit releases it
    const Geometry old = *g;
    Deferred d = take_geometry(*g);
    r = install(ctx, *g, b, true);
    if (r != PYR_OK) {
        *g = old;  // nothing of the old version was released
        pyrhip_built_free(&b, ctx->stream);
        return r;
    }
    ctx->deferred.push_back(d);
    return PYR_OK;
}

// geometryCreate (context.zig:799-816): a host DAG uploaded into the pools
PYR_API PyrResult pyr_geometry_create(PyrContext* ctx, const PyrDag* dag, PyrGeometry* out) {
    diag_clear();
    if (!ctx || !dag || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->nodes.empty() || dag->root >= dag->nodes.size()) return fail(PYR_ERROR_INVALID_ARGUMENT, "empty DAG");
    uint32_t slot = free_geometry_slot(ctx);
    if (slot >= ctx->max_geometries) return fail(PYR_ERROR_CAPACITY, "max_geometries reached");
    const uint64_t nw = dag->nodes.size(), nl = dag->leaves.size(), na = dag->has_attrs ? dag->attrs.size() : 0;
    const int64_t no = ctx->nodes.ra.alloc(nw), lo = ctx->leaves.ra.alloc(nl), ao = ctx->attrs.ra.alloc(na);
    auto release = [&]() {
        if (no >= 0) ctx->nodes.ra.release((uint64_t)no, nw);
        if (lo >= 0) ctx->leaves.ra.release((uint64_t)lo, nl);
        if (ao >= 0) ctx->attrs.ra.release((uint64_t)ao, na);
    };
    if (no < 0 || lo < 0 || ao < 0) {
        release();
        return fail(PYR_ERROR_OUT_OF_MEMORY, "geometry pool full (%llu words, %llu leaves, %llu attributes)",
                    (unsigned long long)nw, (unsigned long long)nl, (unsigned long long)na);
    }
    hipStream_t s = ctx->stream;
    PyrResult r = hip_check(hipMemcpyAsync((char*)ctx->nodes.base + (uint64_t)no * 4, dag->nodes.data(), nw * 4,
                                           hipMemcpyHostToDevice, s),
                            "node upload");
    if (r == PYR_OK && nl)
        r = hip_check(hipMemcpyAsync((char*)ctx->leaves.base + (uint64_t)lo * 8, dag->leaves.data(), nl * 8,
                                     hipMemcpyHostToDevice, s),
                      "leaf upload");
    if (r == PYR_OK && na)
        r = hip_check(hipMemcpyAsync((char*)ctx->attrs.base + (uint64_t)ao * 4, dag->attrs.data(), na * 4,
                                     hipMemcpyHostToDevice, s),
                      "attribute upload");
    // the host arrays may be freed when this returns
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(s), "geometry upload");
    if (r != PYR_OK) {
        release();
        return r;
    }
    if (slot == ctx->geos.size()) ctx->geos.emplace_back();
    Geometry& g = ctx->geos[slot];
    const uint32_t gen = g.gen + 1;
    g = Geometry();
    g.gen = gen;
    g.data.node_offset = (uint32_t)no;
    g.data.leaf_offset = (uint32_t)lo;
    g.data.attribute_offset = (uint32_t)ao;
    g.data.root = dag->root;
    g.data.log2_size = dag->log2_size;
    g.data.flags = dag->has_attrs ? PYR_GEOMETRY_HAS_ATTRIBUTES : 0;
    g.data.default_attribute = 1;
    g.node_words = nw;
    g.leaf_count = nl;
    g.attr_count = na;
    g.voxel_count = (uint32_t)dag->voxel_count;
    g.live = true;
    *out = (PyrGeometry)(uintptr_t)make_handle(slot, g.gen);
    return PYR_OK;
}

// geometryDownsample (context.zig:1036-1050): a coarser LOD built on the GPU
// from the kept sorted voxels of an editable geometry
PYR_API PyrResult pyr_geometry_downsample(PyrContext* ctx, PyrGeometry geometry, uint32_t shift, uint32_t flags,
                                          PyrGeometry* out) {
    diag_clear();
    if (!ctx || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    Geometry* src = geo_of(ctx, geometry);
    if (!src) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    if (!src->keys) return fail(PYR_ERROR_INVALID_ARGUMENT, "source is not editable (build it with PYR_BUILD_EDITABLE)");
    uint32_t slot = free_geometry_slot(ctx);
    if (slot >= ctx->max_geometries) return fail(PYR_ERROR_CAPACITY, "max_geometries reached");
    PyrhipBuilt b{};
    PyrResult r = pyrhip_downsample((uint64_t)(uintptr_t)src->keys, (uint64_t)(uintptr_t)src->vals, src->voxel_count,
                                    src->data.log2_size, shift, nullptr, ctx->stream, &b);
    if (r != PYR_OK) return r;
    if (slot == ctx->geos.size()) ctx->geos.emplace_back();  // may move the vector: src is not used below
    Geometry& g = ctx->geos[slot];
    const uint32_t gen = g.gen + 1;
    g = Geometry();
    g.gen = gen;
    r = install(ctx, g, b, (flags & PYR_BUILD_EDITABLE) != 0);
    if (r != PYR_OK) {
        pyrhip_built_free(&b, ctx->stream);
        return r;
    }
    g.live = true;
    *out = (PyrGeometry)(uintptr_t)make_handle(slot, g.gen);
    return PYR_OK;
}

// geometryDownload (context.zig:1009-1033): the geometry's pool ranges copied
// into a host DAG; a chunk of a streamed world brings its batch's node and
// leaf ranges (the chunk's root refers into them) and its own attributes
PYR_API PyrResult pyr_geometry_download(PyrContext* ctx, PyrGeometry geometry, PyrDag** out) {
    diag_clear();
    if (!ctx || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    Geometry* g = geo_of(ctx, geometry);
    if (!g) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    uint64_t nw = g->node_words, nl = g->leaf_count;
    if (g->shared >= 0) {
        nw = ctx->shared[(size_t)g->shared].node_words;
        nl = ctx->shared[(size_t)g->shared].leaf_count;
    }
    const bool has_attrs = (g->data.flags & PYR_GEOMETRY_HAS_ATTRIBUTES) != 0;
    const uint64_t na = has_attrs ? (g->shared >= 0 ? g->voxel_count : g->attr_count) : 0;
    PyrDag* d = new (std::nothrow) PyrDag();
    if (!d) return fail(PYR_ERROR_OUT_OF_MEMORY, "host DAG");
    d->log2_size = g->data.log2_size;
    d->root = g->data.root;
    d->has_attrs = has_attrs;
    d->nodes.resize(nw);
    d->leaves.resize(nl);
    d->attrs.resize(na);
    hipStream_t s = ctx->stream;
    PyrResult r = hip_check(hipMemcpyAsync(d->nodes.data(), (char*)ctx->nodes.base + (uint64_t)g->data.node_offset * 4,
                                           nw * 4, hipMemcpyDeviceToHost, s),
                            "node download");
    if (r == PYR_OK && nl)
        r = hip_check(hipMemcpyAsync(d->leaves.data(), (char*)ctx->leaves.base + (uint64_t)g->data.leaf_offset * 8,
                                     nl * 8, hipMemcpyDeviceToHost, s),
                      "leaf download");
    const bool palette = (g->data.flags & PYR_GEOMETRY_PALETTE) != 0;
    std::vector<uint32_t> pal(16), packed(palette ? (na + 7) / 8 : 0);
    if (r == PYR_OK && na && !palette)
        r = hip_check(hipMemcpyAsync(d->attrs.data(), (char*)ctx->attrs.base + (uint64_t)g->data.attribute_offset * 4,
                                     na * 4, hipMemcpyDeviceToHost, s),
                      "attribute download");
    if (r == PYR_OK && na && palette) {
        r = hip_check(hipMemcpyAsync(pal.data(), (char*)ctx->attrs.base + (uint64_t)g->data.palette_offset * 4, 64,
                                     hipMemcpyDeviceToHost, s),
                      "palette download");
        if (r == PYR_OK)
            r = hip_check(hipMemcpyAsync(packed.data(),
                                         (char*)ctx->attrs.base + (uint64_t)g->data.attribute_offset * 4,
                                         packed.size() * 4, hipMemcpyDeviceToHost, s),
                          "attribute download");
    }
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(s), "geometry download");
    // palette chunks come back with their 32-bit attributes (attributeAt, dag.zig:64-68)
    if (r == PYR_OK && palette)
        for (uint64_t i = 0; i < na; i++) d->attrs[i] = pal[(packed[i >> 3] >> ((i & 7) * 4)) & 15];
    if (r == PYR_OK && (uint64_t)d->root + 1 >= nw) r = fail(PYR_ERROR_INVALID_ARGUMENT, "geometry without nodes");
    if (r != PYR_OK) {
        delete d;
        return r;
    }
    d->voxel_count = d->nodes[d->root + 1];  // the root's voxel count word
    *out = d;
    return PYR_OK;
}

PYR_API PyrResult pyr_geometry_destroy(PyrContext* ctx, PyrGeometry geometry) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Geometry* g = geo_of(ctx, geometry);
    if (!g) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    if (g->users) return fail(PYR_ERROR_IN_USE, "geometry still has %u instances", g->users);
    // released by the next pyr_commit: instances destroyed since the last
    // commit still reach it through the committed tables (context.zig:1446-1487)
    Deferred d = take_geometry(*g);
    d.slot = (int32_t)(g - ctx->geos.data());
    g->live = false;
    g->held = true;
    ctx->deferred.push_back(d);
    return PYR_OK;
}

PYR_API uint32_t pyr_geometry_index(PyrGeometry geometry) {
    return geometry ? handle_index((uint64_t)(uintptr_t)geometry) : 0xFFFFFFFFu;
}

PYR_API PyrResult pyr_instance_create(PyrContext* ctx, PyrGeometry geometry, PyrInstance* out) {
    diag_clear();
    if (!ctx || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    Geometry* g = geo_of(ctx, geometry);
    if (!g) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    uint32_t slot = 0;
    while (slot < ctx->insts.size() && ctx->insts[slot].live) slot++;
    if (slot >= ctx->max_instances) return fail(PYR_ERROR_CAPACITY, "max_instances reached");
    if (slot == ctx->insts.size()) ctx->insts.emplace_back();
    Instance& in = ctx->insts[slot];
    uint32_t gen = in.gen + 1;
    in = Instance();
    in.gen = gen;
    in.live = true;
    PyrInstanceData& d = in.data;
    static const float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    memcpy(d.object_to_world, ident, sizeof(ident));
    memcpy(d.world_to_object, ident, sizeof(ident));
    d.geometry = handle_index((uint64_t)(uintptr_t)geometry);
    transformed_box(d.object_to_world, (float)(1u << g->data.log2_size), d.bounds_min, d.bounds_max);
    d.mask = 0xFF;
    d.history = ++ctx->history;
    d.flags = PYR_INSTANCE_ACTIVE;
    g->users++;
    *out = (PyrInstance)(uintptr_t)make_handle(slot, in.gen);
    return PYR_OK;
}

PYR_API PyrResult pyr_instance_destroy(PyrContext* ctx, PyrInstance instance) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    Geometry& g = ctx->geos[in->data.geometry];
    if (g.users) g.users--;
    in->live = false;
    return PYR_OK;
}

PYR_API PyrResult pyr_instance_set_transform(PyrContext* ctx, PyrInstance instance, const float m[12]) {
    diag_clear();
    if (!ctx || !m) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    float inv[12];
    if (!inverse34(m, inv)) return fail(PYR_ERROR_INVALID_ARGUMENT, "singular transform");
    memcpy(in->data.object_to_world, m, 48);
    memcpy(in->data.world_to_object, inv, 48);
    const Geometry& g = ctx->geos[in->data.geometry];
    transformed_box(m, (float)(1u << g.data.log2_size), in->data.bounds_min, in->data.bounds_max);
    return PYR_OK;
}

PYR_API PyrResult pyr_instance_set_mask(PyrContext* ctx, PyrInstance instance, uint32_t mask) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    in->data.mask = mask;
    return PYR_OK;
}

// instanceSetGeometry / instanceSetUser (context.zig:1619-1646)
PYR_API PyrResult pyr_instance_set_geometry(PyrContext* ctx, PyrInstance instance, PyrGeometry geometry) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    Geometry* g = geo_of(ctx, geometry);
    if (!g) return fail(PYR_ERROR_INVALID_HANDLE, "geometry");
    Geometry& old = ctx->geos[in->data.geometry];
    if (old.users) old.users--;
    g->users++;
    in->data.geometry = handle_index((uint64_t)(uintptr_t)geometry);
    in->data.history = ++ctx->history;  // another shape: no history
    transformed_box(in->data.object_to_world, (float)(1u << g->data.log2_size), in->data.bounds_min,
                    in->data.bounds_max);
    return PYR_OK;
}

PYR_API PyrResult pyr_instance_reset_history(PyrContext* ctx, PyrInstance instance) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    in->data.history = ++ctx->history;
    return PYR_OK;
}

PYR_API PyrResult pyr_instance_set_user(PyrContext* ctx, PyrInstance instance, uint32_t user) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    Instance* in = inst_of(ctx, instance);
    if (!in) return fail(PYR_ERROR_INVALID_HANDLE, "instance");
    in->data.user = user;
    return PYR_OK;
}

PYR_API uint32_t pyr_instance_index(PyrInstance instance) {
    return instance ? handle_index((uint64_t)(uintptr_t)instance) : 0xFFFFFFFFu;
}

// commit: upload geometry and instance tables and the scene block; the
// previous frame's instances become instances_prev
PYR_API PyrResult pyr_commit(PyrContext* ctx, const PyrFrameInfo* info) {
    diag_clear();
    if (!ctx) return fail(PYR_ERROR_INVALID_ARGUMENT, "null context");
    hipStream_t s = ctx->stream;
    std::swap(ctx->inst_dev[0], ctx->inst_dev[1]);
    std::vector<PyrGeometryData> geo(ctx->geos.size());
    for (size_t i = 0; i < geo.size(); i++) geo[i] = ctx->geos[i].data;
    std::vector<PyrInstanceData> ins(ctx->insts.size());
    uint32_t levels = 3;
    for (size_t i = 0; i < ins.size(); i++) {
        ins[i] = ctx->insts[i].data;
        if (!ctx->insts[i].live) ins[i].flags = 0;
        if (ins[i].geometry < geo.size()) levels = std::max(levels, geo[ins[i].geometry].log2_size);
    }
    ctx->time_prev = ctx->time;
    memcpy(ctx->origin_prev, ctx->origin, sizeof(ctx->origin));
    if (info) {
        ctx->time = info->time;
        memcpy(ctx->origin, info->origin, sizeof(ctx->origin));
    }
    ctx->frame++;
    PyrScene sc;
    memset(&sc, 0, sizeof(sc));
    sc.nodes = (uint64_t)(uintptr_t)ctx->nodes.base;
    sc.leaves = (uint64_t)(uintptr_t)ctx->leaves.base;
    sc.attributes = (uint64_t)(uintptr_t)ctx->attrs.base;
    sc.geometries = (uint64_t)(uintptr_t)ctx->geo_dev;
    sc.instances = (uint64_t)(uintptr_t)ctx->inst_dev[0];
    sc.instances_prev = (uint64_t)(uintptr_t)ctx->inst_dev[1];
    sc.instance_count = (uint32_t)ins.size();
    sc.geometry_count = (uint32_t)geo.size();
    sc.frame = ctx->frame;
    sc.time = ctx->time;
    sc.time_prev = ctx->time_prev;
    memcpy(sc.origin, ctx->origin, sizeof(sc.origin));
    memcpy(sc.origin_prev, ctx->origin_prev, sizeof(sc.origin_prev));
    // material sets of the traversal (context.zig:766-767, 1948-1958)
    sc.materials = (uint64_t)(uintptr_t)ctx->materials_dev;
    for (uint32_t i = 0; i < (uint32_t)ctx->materials.size(); i++) {
        if (ctx->materials[i].flags & PYR_MATERIAL_TRANSPARENT) sc.transparent_materials[i >> 6] |= 1ull << (i & 63);
        if (ctx->materials[i].flags & PYR_MATERIAL_CUTOUT) sc.cutout_materials[i >> 6] |= 1ull << (i & 63);
    }
    PyrResult r = PYR_OK;
    if (!geo.empty())
        r = hip_check(hipMemcpyAsync(ctx->geo_dev, geo.data(), geo.size() * sizeof(PyrGeometryData),
                                     hipMemcpyHostToDevice, s),
                      "geometry table");
    if (r == PYR_OK && !ins.empty())
        r = hip_check(hipMemcpyAsync(ctx->inst_dev[0], ins.data(), ins.size() * sizeof(PyrInstanceData),
                                     hipMemcpyHostToDevice, s),
                      "instance table");
    if (r == PYR_OK)
        r = hip_check(hipMemcpyAsync(ctx->scene_dev, &sc, sizeof(sc), hipMemcpyHostToDevice, s), "scene");
    // host staging above is stack memory: finish the copies before returning
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(s), "commit");
    // a failed commit may have replaced some tables: keep a stack that fits both scenes
    ctx->scene_levels = r == PYR_OK ? levels : std::max(ctx->scene_levels, levels);
    const bool tr = (sc.transparent_materials[0] | sc.transparent_materials[1] | sc.transparent_materials[2] |
                     sc.transparent_materials[3]) != 0;
    const bool cu = (sc.cutout_materials[0] | sc.cutout_materials[1] | sc.cutout_materials[2] |
                     sc.cutout_materials[3]) != 0;
    ctx->scene_transparent = r == PYR_OK ? tr : (ctx->scene_transparent || tr);
    ctx->scene_cutout = r == PYR_OK ? cu : (ctx->scene_cutout || cu);
    // the new tables are in place and the work queued before them is done:
    // nothing reaches the deferred ranges any more
    if (r == PYR_OK) {
        for (const Deferred& d : ctx->deferred) release_deferred(ctx, d);
        ctx->deferred.clear();
    }
    return r;
}

PYR_API uint64_t pyr_scene_device(PyrContext* ctx) { return ctx ? (uint64_t)(uintptr_t)ctx->scene_dev : 0; }

// trace (context.zig:1869-1885); ray_mask 0 = all instances; PYR_TRACE_EXTENDED
// writes PyrHitEx (gpu_kernels.zig:74-85)
PYR_API PyrResult pyr_trace(PyrContext* ctx, uint64_t rays, uint64_t hits, uint32_t count, uint32_t ray_mask,
                            uint32_t flags) {
    diag_clear();
    if (!ctx || (count && (!rays || !hits))) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    const bool skipping = ((flags & PYR_TRACE_SKIP_TRANSPARENT) && ctx->scene_transparent) || ctx->scene_cutout;
    PyrResult r = launch_trace_scene(ctx->scene_dev, rays, hits, count, ray_mask ? ray_mask : 0xFFFFFFFFu, flags,
                                     ctx->scene_levels, skipping, ctx->stream);
    if (r == PYR_OK && (ctx->flags & PYR_CREATE_DEBUG)) r = hip_check(hipStreamSynchronize(ctx->stream), "trace");
    return r;
}

PYR_API PyrResult pyr_view_create(PyrContext* ctx, PyrView* out) {
    diag_clear();
    if (!ctx || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    *out = (PyrView)(uintptr_t)make_handle(ctx->views++, 1);
    return PYR_OK;
}

PYR_API PyrResult pyr_view_destroy(PyrContext* ctx, PyrView view) {
    diag_clear();
    if (!ctx || !view) return fail(PYR_ERROR_INVALID_HANDLE, "view");
    return PYR_OK;
}

PYR_API PyrResult pyr_view_reset_history(PyrContext* ctx, PyrView view) {
    diag_clear();
    if (!ctx || !view) return fail(PYR_ERROR_INVALID_HANDLE, "view");
    return PYR_OK;
}

// jitter sequences (root.zig:405-443)
namespace {
// radical inverse of index in base b: its base-b digits mirrored behind the point
float radical_inverse(uint32_t index, uint32_t b) {
    float scale = 1.0f, r = 0.0f;
    for (uint32_t i = index; i > 0; i /= b) {
        scale /= (float)b;
        r += scale * (float)(i % b);
    }
    return r;
}
}  // namespace

PYR_API void pyr_jitter_halton(uint32_t frame, float out[2]) {
    if (!out) return;
    const uint32_t i = frame % 1024u + 1u;
    out[0] = radical_inverse(i, 2) - 0.5f;
    out[1] = radical_inverse(i, 3) - 0.5f;
}

PYR_API void pyr_jitter_ordered(uint32_t frame, float out[2]) {
    if (!out) return;
    const uint32_t k = frame % 16u;
    const uint32_t row = ((k & 1u) << 1) | ((k >> 1) & 1u);  // 0, 2, 1, 3
    const uint32_t step = (k >> 2) & 3u;
    const uint32_t col = (row & 1u) ? 3u - step : step;
    out[0] = ((float)col + 0.5f) / 4.0f - 0.5f;
    out[1] = ((float)row + 0.5f) / 4.0f - 0.5f;
}

// materials (root.zig:363-370, context.zig:1902-1926, 1960-1964)
PYR_API void pyr_material_default(PyrMaterial* m) {
    if (!m) return;
    memset(m, 0, sizeof(*m));
    for (int a = 0; a < 3; a++) {
        m->base_color[a] = 1.0f;
        m->subsurface_color[a] = 1.0f;
    }
    m->roughness = 0.8f;
    m->flags = PYR_MATERIAL_VOXEL_COLOR;
    m->ior = 1.5f;
    m->wave_height = 0.15f;
    m->wave_length = 12.0f;
    m->wave_speed = 0.35f;
    m->clearcoat_roughness = 0.1f;
    m->texture_scale = 1.0f;
    m->normal_scale = 1.0f;
}

PYR_API PyrResult pyr_material_set(PyrContext* ctx, uint32_t index, const PyrMaterial* m) {
    diag_clear();
    if (!ctx || !m) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (index >= PYR_MAX_MATERIALS)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "material index %u >= %u", index, PYR_MAX_MATERIALS);
    ctx->materials[index] = *m;
    // from the host copy, which stays put: the copy may finish after the return
    return hip_check(hipMemcpyAsync(ctx->materials_dev + index, &ctx->materials[index], sizeof(PyrMaterial),
                                    hipMemcpyHostToDevice, ctx->stream),
                     "material upload");
}

PYR_API uint32_t pyr_voxel_attribute(uint32_t material, uint32_t r, uint32_t g, uint32_t b) {
    return PYR_VOXEL(material, r, g, b);
}

// render (context.zig:1766-1867), primary hits and linear depth only
PYR_API PyrResult pyr_render(PyrContext* ctx, PyrView view, const PyrCamera* camera, const PyrTargets* t) {
    diag_clear();
    if (!ctx || !camera || !t) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (!view) return fail(PYR_ERROR_INVALID_HANDLE, "view");
    if (t->motion || t->color || t->normal || t->albedo || t->material)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "only hits and depth targets are produced by this backend");
    if (camera->width == 0 || camera->height == 0) return PYR_OK;
    float w2v[12];
    if (!inverse34(camera->view_to_world, w2v)) return fail(PYR_ERROR_INVALID_ARGUMENT, "singular camera pose");
    PyrResult r = hip_check(hipMemcpyAsync(ctx->w2v_dev, w2v, sizeof(w2v), hipMemcpyHostToDevice, ctx->stream),
                            "camera upload");
    if (r == PYR_OK)
        r = launch_render_scene(ctx->scene_dev, *camera, ctx->w2v_dev, t->hits, t->depth,
                                t->ray_mask ? t->ray_mask : 0xFFFFFFFFu, t->flags & PYR_TRACE_NO_ATTRIBUTE,
                                ctx->scene_levels, ctx->scene_cutout, ctx->stream);
    // w2v is host stack memory
    if (r == PYR_OK) r = hip_check(hipStreamSynchronize(ctx->stream), "render");
    return r;
}

// camera helpers (root.zig:449-481)
static void normalize3(const float v[3], float r[3]) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    r[0] = v[0] / l;
    r[1] = v[1] / l;
    r[2] = v[2] / l;
}

static void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

PYR_API void pyr_camera_look_at(PyrCamera* c, const float eye[3], const float target[3], const float up[3]) {
    if (!c || !eye || !target || !up) return;
    const float b0[3] = {eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]};
    float back[3], r0[3], right[3], tu[3];
    normalize3(b0, back);
    cross3(up, back, r0);
    normalize3(r0, right);
    cross3(back, right, tu);
    const float m[12] = {right[0], tu[0], back[0], eye[0], right[1], tu[1], back[1], eye[1],
                         right[2], tu[2], back[2], eye[2]};
    memcpy(c->view_to_world, m, sizeof(m));
}

// root.zig:483-489
PYR_API void pyr_camera_orthographic(PyrCamera* c, float half_height, uint32_t width, uint32_t height) {
    if (!c) return;
    c->projection = PYR_PROJECTION_ORTHOGRAPHIC;
    c->width = width;
    c->height = height;
    c->scale[0] = half_height * (float)width / (float)height;
    c->scale[1] = half_height;
}

PYR_API void pyr_camera_perspective(PyrCamera* c, float fov_y, uint32_t width, uint32_t height, float near_plane) {
    if (!c) return;
    const float ty = tanf(fov_y * 0.5f);
    c->projection = PYR_PROJECTION_PERSPECTIVE;
    c->width = width;
    c->height = height;
    c->scale[0] = ty * (float)width / (float)height;
    c->scale[1] = ty;
    c->near_plane = near_plane;
}

}  // extern "C"
