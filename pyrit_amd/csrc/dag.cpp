// dag.cpp - host DAGs (pyr_dag_*, the reference's src/dag_builder.zig and
// src/io.zig entry points as exported by src/root.zig): voxel lists, dense
// grids and voxel functions are built on the GPU by pyrhip_build and the
// arrays are downloaded; the binary "PYRD" format is the reference's.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

// GPU build of host voxel records into a host DAG (each call on its own stream)
PyrResult build_host(uint32_t log2_size, const PyrVoxel* voxels, size_t count, uint32_t flags, PyrDag** out) {
    if (!out || (count && !voxels)) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    if (count >= 0xFFFFFFFFull) return fail(PYR_ERROR_INVALID_ARGUMENT, "too many voxels (%zu)", count);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PYR_ERROR_CUDA, "no HIP device for the DAG build");
    }
    void* dev = nullptr;
    PyrResult r = PYR_OK;
    if (pool_alloc(&dev, std::max<size_t>(count, 1) * sizeof(PyrVoxel), s) != hipSuccess) {
        (void)hipGetLastError();
        r = fail(PYR_ERROR_OUT_OF_MEMORY, "DAG build input of %zu voxels", count);
    }
    if (r == PYR_OK && count &&
        hipMemcpyAsync(dev, voxels, count * sizeof(PyrVoxel), hipMemcpyHostToDevice, s) != hipSuccess)
        r = fail(PYR_ERROR_CUDA, "voxel upload");
    PyrhipBuilt b{};
    if (r == PYR_OK) r = pyrhip_build(log2_size, 0, 0, 0, (uint64_t)(uintptr_t)dev, (uint32_t)count, nullptr, s, &b);
    PyrDag* d = nullptr;
    if (r == PYR_OK) {
        d = new (std::nothrow) PyrDag();
        if (!d) r = fail(PYR_ERROR_OUT_OF_MEMORY, "host DAG");
    }
    if (r == PYR_OK) {
        d->log2_size = log2_size;
        d->root = b.root;
        d->voxel_count = b.voxel_count;
        d->has_attrs = !(flags & PYR_DAG_NO_ATTRIBUTES);
        d->nodes.resize(b.node_words);
        d->leaves.resize(b.leaf_count);
        if (d->has_attrs) d->attrs.resize(b.voxel_count);
        bool ok = hipMemcpyAsync(d->nodes.data(), (void*)(uintptr_t)b.nodes, d->nodes.size() * 4,
                                 hipMemcpyDeviceToHost, s) == hipSuccess;
        if (ok && !d->leaves.empty())
            ok = hipMemcpyAsync(d->leaves.data(), (void*)(uintptr_t)b.leaves, d->leaves.size() * 8,
                                hipMemcpyDeviceToHost, s) == hipSuccess;
        if (ok && !d->attrs.empty())
            ok = hipMemcpyAsync(d->attrs.data(), (void*)(uintptr_t)b.attrs, d->attrs.size() * 4,
                                hipMemcpyDeviceToHost, s) == hipSuccess;
        if (ok) ok = hipStreamSynchronize(s) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            delete d;
            d = nullptr;
            r = fail(PYR_ERROR_CUDA, "DAG download");
        }
    }
    pyrhip_built_free(&b, s);
    pool_free(dev, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (r == PYR_OK) *out = d;
    return r;
}

// region descent of buildFn (dag_builder.zig): skip cubes the caller reports
// empty, sample the voxel function inside the 4^3 bricks that remain
void collect_fn(PyrVoxelFn voxel, PyrRegionEmptyFn region_empty, void* user, int32_t x, int32_t y, int32_t z,
                uint32_t size, std::vector<PyrVoxel>& out) {
    if (region_empty && region_empty(user, x, y, z, size)) return;
    if (size <= 4) {
        for (uint32_t dz = 0; dz < size; dz++)
            for (uint32_t dy = 0; dy < size; dy++)
                for (uint32_t dx = 0; dx < size; dx++) {
                    const uint32_t a = voxel(user, x + (int32_t)dx, y + (int32_t)dy, z + (int32_t)dz);
                    if (a) out.push_back(PyrVoxel{x + (int32_t)dx, y + (int32_t)dy, z + (int32_t)dz, a});
                }
        return;
    }
    const uint32_t h = size / 2;
    for (uint32_t c = 0; c < 8; c++)
        collect_fn(voxel, region_empty, user, x + (int32_t)((c & 1) * h), y + (int32_t)(((c >> 1) & 1) * h),
                   z + (int32_t)((c >> 2) * h), h, out);
}

constexpr uint32_t kPyrdVersion = 1;
constexpr size_t kPyrdHeader = 4 + 4 * 4 + 8 * 4;

template <class T>
void put(uint8_t*& p, T v) {
    memcpy(p, &v, sizeof(T));  // little-endian host
    p += sizeof(T);
}
template <class T>
T get(const uint8_t*& p) {
    T v;
    memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
}

}  // namespace
}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyr_dag_build_points(uint32_t log2_size, const PyrVoxel* voxels, size_t count, uint32_t flags,
                                       PyrDag** out) {
    diag_clear();
    if (!out || (count && !voxels)) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    // buildPoints (dag_builder.zig:333-335): every point inside the cube, no empty attribute
    const int64_t n = 1ll << log2_size;
    for (size_t i = 0; i < count; i++) {
        const PyrVoxel& v = voxels[i];
        if (v.x < 0 || v.y < 0 || v.z < 0 || v.x >= n || v.y >= n || v.z >= n || v.attribute == 0)
            return fail(PYR_ERROR_INVALID_ARGUMENT, "voxel %zu (%d, %d, %d, attribute %u) outside [0, %lld)^3 or empty",
                        i, v.x, v.y, v.z, v.attribute, (long long)n);
    }
    return build_host(log2_size, voxels, count, flags, out);
}

PYR_API PyrResult pyr_dag_build_dense(uint32_t log2_size, const uint32_t* voxels, uint32_t flags, PyrDag** out) {
    diag_clear();
    if (!voxels || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 10)
        return fail(PYR_ERROR_INVALID_ARGUMENT, "dense grids need log2_size in [3, 10]");
    const uint64_t n = 1ull << log2_size;
    std::vector<PyrVoxel> pts;
    for (uint64_t z = 0; z < n; z++)
        for (uint64_t y = 0; y < n; y++)
            for (uint64_t x = 0; x < n; x++) {
                const uint32_t a = voxels[x + n * (y + n * z)];
                if (a) pts.push_back(PyrVoxel{(int32_t)x, (int32_t)y, (int32_t)z, a});
            }
    return build_host(log2_size, pts.data(), pts.size(), flags, out);
}

PYR_API PyrResult pyr_dag_build_fn(uint32_t log2_size, PyrVoxelFn voxel, PyrRegionEmptyFn region_empty, void* user,
                                   uint32_t flags, PyrDag** out) {
    diag_clear();
    if (!voxel || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (log2_size < 3 || log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size must be in [3, 20]");
    std::vector<PyrVoxel> pts;
    collect_fn(voxel, region_empty, user, 0, 0, 0, 1u << log2_size, pts);
    return build_host(log2_size, pts.data(), pts.size(), flags, out);
}

PYR_API void pyr_dag_get_info(const PyrDag* d, PyrDagInfo* out) {
    if (!d || !out) return;
    out->log2_size = d->log2_size;
    out->root = d->root;
    out->voxel_count = d->voxel_count;
    out->nodes = d->nodes.data();
    out->node_words = d->nodes.size();
    // empty arrays still get a valid pointer; attributes is NULL only for a
    // geometry-only DAG
    static const uint64_t kNoLeaf = 0;
    static const uint32_t kNoAttr = 0;
    out->leaves = d->leaves.empty() ? &kNoLeaf : d->leaves.data();
    out->leaf_count = d->leaves.size();
    out->attributes = !d->has_attrs ? nullptr : d->attrs.empty() ? &kNoAttr : d->attrs.data();
    out->attribute_count = d->has_attrs ? d->attrs.size() : 0;
}

PYR_API void pyr_dag_destroy(PyrDag* d) { delete d; }

// "PYRD", version, log2_size, root, flags (bit 0: attributes), voxel_count,
// node_words, leaf_count, attribute count, nodes, leaves, attributes; all
// little-endian (src/io.zig:86-125)
PYR_API PyrResult pyr_dag_save(const PyrDag* d, void* buffer, size_t* size) {
    diag_clear();
    if (!d || !size) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    const uint64_t na = d->has_attrs ? d->attrs.size() : 0;
    const size_t need = kPyrdHeader + d->nodes.size() * 4 + d->leaves.size() * 8 + na * 4;
    if (!buffer) {
        *size = need;
        return PYR_OK;
    }
    if (*size < need) return fail(PYR_ERROR_INVALID_ARGUMENT, "buffer of %zu bytes, %zu needed", *size, need);
    uint8_t* p = (uint8_t*)buffer;
    memcpy(p, "PYRD", 4);
    p += 4;
    put<uint32_t>(p, kPyrdVersion);
    put<uint32_t>(p, d->log2_size);
    put<uint32_t>(p, d->root);
    put<uint32_t>(p, d->has_attrs ? 1u : 0u);
    put<uint64_t>(p, d->voxel_count);
    put<uint64_t>(p, d->nodes.size());
    put<uint64_t>(p, d->leaves.size());
    put<uint64_t>(p, na);
    memcpy(p, d->nodes.data(), d->nodes.size() * 4);
    p += d->nodes.size() * 4;
    if (!d->leaves.empty()) memcpy(p, d->leaves.data(), d->leaves.size() * 8);
    p += d->leaves.size() * 8;
    if (na) memcpy(p, d->attrs.data(), na * 4);
    *size = need;
    return PYR_OK;
}

// io.zig:127-160: the same validation (magic, version, edge, sizes, root,
// attribute count = voxel count)
PYR_API PyrResult pyr_dag_load(const void* data, size_t size, PyrDag** out) {
    diag_clear();
    if (!data || !out) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    const uint8_t* p = (const uint8_t*)data;
    if (size < kPyrdHeader || memcmp(p, "PYRD", 4) != 0) return fail(PYR_ERROR_INVALID_ARGUMENT, "not a PYRD DAG");
    p += 4;
    const uint32_t version = get<uint32_t>(p), log2 = get<uint32_t>(p), root = get<uint32_t>(p);
    const bool has_attrs = (get<uint32_t>(p) & 1) != 0;
    const uint64_t voxels = get<uint64_t>(p), nw = get<uint64_t>(p), nl = get<uint64_t>(p), na = get<uint64_t>(p);
    if (version != kPyrdVersion) return fail(PYR_ERROR_INVALID_ARGUMENT, "PYRD version %u", version);
    if (log2 < 3 || log2 > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "PYRD edge 2^%u", log2);
    const uint64_t lim = ~0ull / 16;
    if (nw > lim || nl > lim || na > lim) return fail(PYR_ERROR_INVALID_ARGUMENT, "PYRD sizes");
    const uint64_t need = kPyrdHeader + nw * 4 + nl * 8 + na * 4;
    if (need > size || root >= nw || (has_attrs && na != voxels))
        return fail(PYR_ERROR_INVALID_ARGUMENT, "PYRD data inconsistent or truncated");
    PyrDag* d = new (std::nothrow) PyrDag();
    if (!d) return fail(PYR_ERROR_OUT_OF_MEMORY, "host DAG");
    d->log2_size = log2;
    d->root = root;
    d->voxel_count = voxels;
    d->has_attrs = has_attrs;
    d->nodes.resize(nw);
    d->leaves.resize(nl);
    memcpy(d->nodes.data(), p, nw * 4);
    p += nw * 4;
    if (nl) memcpy(d->leaves.data(), p, nl * 8);
    p += nl * 8;
    if (has_attrs) {
        d->attrs.resize(na);
        if (na) memcpy(d->attrs.data(), p, na * 4);
    }
    *out = d;
    return PYR_OK;
}

}  // extern "C"
