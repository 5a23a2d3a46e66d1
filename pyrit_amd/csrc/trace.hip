// trace.hip - software sparse-voxel-DAG ray traversal for gfx950.
//
// One ray per lane. The traversal is the reference's walk (src/device/dag.zig:
// 160-351: octant mirroring, ESVO-style ascent through the highest changed
// coordinate bit, the two brick levels inside the 64-bit leaf) and must give
// the same hits and the same fp32 t bit for bit, so this file is compiled with
// -ffp-contract=off and every float expression keeps the reference's order of
// operations. The per-ray node stack (one word per level) lives in LDS,
// strided by lane so the 64 lanes of a wave hit distinct banks, and is sized
// by the launch to the DAG's edge (or the scene's largest), not to the
// largest edge the format allows, which keeps more waves resident.
//
// A hit's attribute needs its DFS rank (a second walk down the stack, one
// dependent load per earlier sibling). The walk computes it after the step
// loop, once the whole wave has left the loop, so all lanes of a wave that
// hit rank their voxels together; inside the loop every lane that hits in a
// different iteration would make the wave run the rank walk again with
// only that lane active.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include "common.h"
#include "internal.h"

namespace pyr {

constexpr int kTraceBlock = 128;  // rays per workgroup (types.zig:498 trace_block)
constexpr int kMaxLevels = 21;    // log2_size <= 20
constexpr uint32_t kMaxSteps = 65536;

struct DagView {
    const uint32_t* nodes;
    const uint64_t* leaves;
    const uint32_t* attrs;    // null: default_attribute
    const uint32_t* palette;  // non-null: attrs holds 4-bit indices into these 16 values
    uint32_t root, log2_size, default_attribute, attr_base;
};

// attributeAt (dag.zig:64-68): rank r's value, or its palette entry through
// the 4-bit index in bits 4 (r & 7) of word r >> 3
__device__ __forceinline__ uint32_t attribute_at(const DagView& g, uint32_t rank) {
    const uint32_t r = g.attr_base + rank;
    if (g.palette) return g.palette[(g.attrs[r >> 3] >> ((r & 7) * 4)) & 15u];
    return g.attrs[r];
}

struct DagHitD {
    float t;
    uint32_t face;
    uint32_t attribute;
    uint32_t inside;
    int32_t voxel[3];
};

__device__ __forceinline__ uint32_t bit_of(uint32_t v, uint32_t s) { return (v >> s) & 1u; }
__device__ __forceinline__ uint32_t brick_bit(uint32_t x, uint32_t y, uint32_t z) { return x + 4 * y + 16 * z; }

// attributeRank (dag.zig:70-87); stk[s * kTraceBlock] is the node whose
// children have edge 2^s
__device__ uint32_t attribute_rank(const DagView& g, const uint32_t* stk, const uint32_t v[3], uint64_t brick) {
    uint32_t rank = 0;
    for (uint32_t s = g.log2_size - 1; s >= 2; s--) {
        const uint32_t node = stk[s * kTraceBlock];
        const uint32_t mask = g.nodes[node] & 0xFFu;
        const uint32_t idx = bit_of(v[0], s) | (bit_of(v[1], s) << 1) | (bit_of(v[2], s) << 2);
        const uint32_t before = (uint32_t)__popc(mask & ((1u << idx) - 1));
        for (uint32_t k = 0; k < before; k++) {
            const uint32_t ref = g.nodes[node + 2 + k];
            rank += s == 2 ? (uint32_t)__popcll(g.leaves[ref]) : g.nodes[ref + 1];
        }
    }
    const uint32_t bit = brick_bit(v[0] & 3, v[1] & 3, v[2] & 3);
    return rank + (uint32_t)__popcll(brick & ((1ull << bit) - 1));
}

// skipped (dag.zig:153-157): bit (attribute & 0xFF) of a 256-bit material set
__device__ __forceinline__ bool skipped(const uint64_t* set, uint32_t attribute) {
    if (!set) return false;
    const uint32_t idx = attribute & 0xFFu;
    return ((set[idx >> 6] >> (idx & 63u)) & 1ull) != 0;
}

// cutHole (dag.zig:133-151) for a DAG at the origin of its object space: is
// the entry point p on face `axis` of the voxel in a hole of the material's
// fixed 3 x 3 pattern (about 30 % of the face)?
__device__ __forceinline__ bool cut_hole(const uint64_t* cut, uint32_t attribute, const int32_t voxel[3],
                                         const float p[3], uint32_t axis) {
    if (!skipped(cut, attribute)) return false;
    uint32_t sub[2] = {0, 0}, k = 0;
#pragma unroll
    for (uint32_t a = 0; a < 3; a++) {
        if (a == axis) continue;
        const float f = fminf(fmaxf(p[a] - (float)voxel[a], 0.0f), 0.999f);
        sub[k++] = (uint32_t)(f * 3.0f);
    }
    uint32_t h = (uint32_t)voxel[0] * 0x8da6b343u + (uint32_t)voxel[1] * 0xd8163841u + (uint32_t)voxel[2] * 0xcb1ab31fu +
                 (attribute >> 8) * 0x2545f491u + axis * 0x9e3779b9u + (sub[0] * 3 + sub[1]) * 0x632be5abu;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return (h & 0xFFu) < 77u;
}

// walk (dag.zig:160-351); stk points at this lane's column of the block's
// LDS stack. SKIPPING (traceSkipping, dag.zig:113-115, 236-289): voxels of a
// material in `skip` are passed, and voxels of a material in `cut` where the
// entry point lies in a hole of the material's pattern or the ray starts
// inside them; such a walk needs each candidate voxel's attribute inside the
// loop, so it ranks there, and its hits carry the attribute even when
// want_attr is false (as the reference's do).
template <bool SKIPPING>
__device__ bool walk(const DagView& g, const float o[3], const float d[3], float tmin, float tmax, bool want_attr,
                     uint32_t* stk, DagHitD* hit, const uint64_t* skip = nullptr, const uint64_t* cut = nullptr) {
    const uint32_t n = 1u << g.log2_size;
    const float nf = (float)n;
    uint32_t oct = 0;
    float od[3], dd[3], inv[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float oa = o[a], da = d[a];
        if (da < 0) {
            oa = nf - oa;
            da = -da;
            oct |= 1u << a;
        }
        od[a] = oa;
        dd[a] = da;
        inv[a] = 1.0f / fmaxf(da, 1e-30f);
    }
    uint32_t axis = 0;
    float tn = -od[0] * inv[0];
#pragma unroll
    for (int a = 1; a < 3; a++) {
        const float ta = -od[a] * inv[a];
        if (ta > tn) {
            tn = ta;
            axis = (uint32_t)a;
        }
    }
    const float tf = fminf(fminf((nf - od[0]) * inv[0], (nf - od[1]) * inv[1]), (nf - od[2]) * inv[2]);
    float t = fmaxf(tn, tmin);
    const float tend = fminf(tf, tmax);
    if (!(t < tend)) return false;
    const bool starts_inside = tn < tmin;
    // brick bit of a mirrored cell: its low two bits per axis, flipped on the
    // mirrored axes (3 - x for a voxel, 2 - x for a 2 x 2 x 2 block; n is a
    // multiple of 4)
    const uint32_t flip1 = (oct & 1u ? 3u : 0u) | (oct & 2u ? 12u : 0u) | (oct & 4u ? 48u : 0u);
    const uint32_t flip2 = flip1 & 0x2Au;
    bool moved = false;

    uint32_t scale = g.log2_size - 1;
    uint32_t cell[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float p = fminf(fmaxf(floorf(od[a] + dd[a] * t), 0.0f), nf - 1.0f);
        cell[a] = (uint32_t)p & ~((1u << scale) - 1);
    }
    stk[scale * kTraceBlock] = g.root;
    uint64_t brick = 0;
    bool found = false;

    // One iteration per cell: look the cell up; a filled leaf voxel ends the
    // walk, a filled larger cell is entered (descend), an empty one is left
    // along the ray (advance, ascending to the highest changed coordinate
    // bit). Both successors are formed in straight-line code and the lookup's
    // result picks one, so a wave whose lanes split between entering and
    // leaving cells runs one short path instead of two branches.
    for (uint32_t step = 0; step < kMaxSteps; step++) {
        bool exists;
        uint32_t child = 0;
        if (scale >= 2) {
            const uint32_t node = stk[scale * kTraceBlock];
            const uint32_t mask = g.nodes[node] & 0xFFu;
            const uint32_t idx = (__builtin_amdgcn_ubfe(cell[0], scale, 1) | (__builtin_amdgcn_ubfe(cell[1], scale, 1) << 1) |
                                  (__builtin_amdgcn_ubfe(cell[2], scale, 1) << 2)) ^ oct;
            const uint32_t bit = 1u << idx;
            exists = (mask & bit) != 0;
            if (exists) child = g.nodes[node + 2 + __popc(mask & (bit - 1))];
        } else {
            const uint32_t b = brick_bit(cell[0] & 3u, cell[1] & 3u, cell[2] & 3u) ^ (scale ? flip2 : flip1);
            // scale 1: any voxel of the 2 x 2 x 2 block
            const uint64_t sel = scale ? 0x00330033ull : 1ull;
            exists = ((brick >> b) & sel) != 0;
        }
        if (exists && scale == 0) {
            if constexpr (SKIPPING) {
                uint32_t v[3];
                int32_t vox[3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    v[q] = ((oct >> q) & 1u) ? n - 1 - cell[q] : cell[q];
                    vox[q] = (int32_t)v[q];
                }
                uint32_t attribute = g.default_attribute;
                if ((want_attr || skip || cut) && g.attrs) attribute = attribute_at(g, attribute_rank(g, stk, v, brick));
                bool through = false;
                if (starts_inside && !moved) {
                    through = skipped(cut, attribute);
                } else if (cut) {
                    float p[3];
#pragma unroll
                    for (int q = 0; q < 3; q++) p[q] = o[q] + d[q] * t;
                    through = cut_hole(cut, attribute, vox, p, axis);
                }
                // a passed voxel is left like an empty one
                if (skipped(skip, attribute) || through) {
                    exists = false;
                } else {
                    hit->attribute = attribute;
                    found = true;
                    break;
                }
            } else {
                found = true;
                break;
            }
        }
        if (exists) {
            if (scale >= 3) stk[(scale - 1) * kTraceBlock] = child;
            else if (scale == 2) brick = g.leaves[child];
        }
        // leaving the cell: the nearest of its three far faces
        const uint32_t size = 1u << scale;
        const uint32_t n0 = cell[0] + size, n1 = cell[1] + size, n2 = cell[2] + size;
        const float t0 = ((float)n0 - od[0]) * inv[0];
        const float t1 = ((float)n1 - od[1]) * inv[1];
        const float t2 = ((float)n2 - od[2]) * inv[2];
        // the first axis with the smallest exit (ties to the lower axis)
        const bool a1 = t1 < t0;
        const float t01 = a1 ? t1 : t0;
        const bool s2 = t2 < t01;
        const bool s1 = a1 && !s2, s0 = !a1 && !s2;
        const float tnext = s2 ? t2 : t01;
        const uint32_t ax = s2 ? 2u : (a1 ? 1u : 0u);
        const uint32_t old = s2 ? cell[2] : (a1 ? cell[1] : cell[0]);
        const uint32_t next = s2 ? n2 : (a1 ? n1 : n0);
        if (!exists && (!(tnext < tend) || next >= n)) break;
        if (exists) {
            // entering it: the child cell that holds the point at t
            const uint32_t half = size >> 1;
            const float halff = (float)half;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float p = od[a] + dd[a] * t;
                if (p - (float)cell[a] >= halff) cell[a] += half;
            }
            scale -= 1;
        } else {
            t = fmaxf(t, tnext);
            moved = true;
            axis = ax;
            cell[0] = s0 ? n0 : cell[0];
            cell[1] = s1 ? n1 : cell[1];
            cell[2] = s2 ? n2 : cell[2];
            const uint32_t h = 31 - (uint32_t)__clz(old ^ next);
            if (h > scale) {
                scale = h;
                const uint32_t m = ~((1u << scale) - 1);
#pragma unroll
                for (int a = 0; a < 3; a++) cell[a] &= m;
            }
        }
    }
    if (!found) return false;
    // the hit: every value below is final when the loop leaves
    uint32_t v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        v[a] = ((oct >> a) & 1u) ? n - 1 - cell[a] : cell[a];
        hit->voxel[a] = (int32_t)v[a];
    }
    const bool mirrored = ((oct >> axis) & 1u) != 0;
    hit->t = t;
    hit->face = 2 * axis + (mirrored ? 0u : 1u);
    hit->inside = starts_inside && !moved;
    if constexpr (!SKIPPING) {
        uint32_t attribute = g.default_attribute;
        // the stack still holds the hit voxel's path and brick is its leaf
        if (want_attr && g.attrs) attribute = attribute_at(g, attribute_rank(g, stk, v, brick));
        hit->attribute = attribute;
    }
    return true;
}

__device__ __forceinline__ PyrHit to_hit(bool found, const DagHitD& h, uint32_t instance) {
    PyrHit r;
    if (found) {
        r.t = h.t;
        r.instance = instance;
        r.attribute = h.attribute;
        r.meta = h.face | (h.inside ? PYR_HIT_INSIDE : 0u);
    } else {
        r.t = PYR_FLT_MAX;
        r.instance = PYR_NO_HIT;
        r.attribute = 0;
        r.meta = 0;
    }
    return r;
}

// ---- scene traversal (trace.zig:56-113) -------------------------------------
__device__ __forceinline__ float safe_rcp(float x) {
    return 1.0f / (fabsf(x) < 1e-30f ? (x < 0 ? -1e-30f : 1e-30f) : x);
}

__device__ __forceinline__ void xform_point(const float* m, const float p[3], float r[3]) {
    r[0] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    r[1] = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    r[2] = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
}

__device__ __forceinline__ void xform_vector(const float* m, const float v[3], float r[3]) {
    r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    r[1] = m[4] * v[0] + m[5] * v[1] + m[6] * v[2];
    r[2] = m[8] * v[0] + m[9] * v[1] + m[10] * v[2];
}

// anyCutout / cutoutHere (trace.zig:161-173): cutouts only where a voxel
// edge is at most about 1.4 world units (row 0 of world_to_object)
__device__ __forceinline__ bool any_cutout(const PyrScene& s) {
    return (s.cutout_materials[0] | s.cutout_materials[1] | s.cutout_materials[2] | s.cutout_materials[3]) != 0;
}
__device__ __forceinline__ bool cutout_here(const PyrInstanceData& in) {
    float w2o = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) w2o += in.world_to_object[a] * in.world_to_object[a];
    return w2o > 0.5f;
}

// SKIPPING: the launch found a material set that may apply (skip flag with
// transparent materials, or cutout materials); each variant holds one walk,
// which keeps the registers of the plain one at its own count. EXT: the best
// hit's point in its instance's object space goes to p_object (trace.zig:107).
template <bool SKIPPING, bool EXT = false>
__device__ PyrHit trace_scene(const PyrScene& s, const float o[3], const float d[3], float tmin, float tmax,
                              uint32_t ray_mask, uint32_t flags, uint32_t* stk, uint32_t levels,
                              float* p_object = nullptr) {
    const PyrInstanceData* inst = (const PyrInstanceData*)(uintptr_t)s.instances;
    const PyrGeometryData* geo = (const PyrGeometryData*)(uintptr_t)s.geometries;
    const float inv[3] = {safe_rcp(d[0]), safe_rcp(d[1]), safe_rcp(d[2])};
    const bool want_attr = (flags & PYR_TRACE_NO_ATTRIBUTE) == 0;
    float best = tmax;
    PyrHit result = to_hit(false, DagHitD{}, 0);
    // material sets of the traversal (trace.zig:80, 93-98)
    const uint64_t* skip = SKIPPING && (flags & PYR_TRACE_SKIP_TRANSPARENT) ? s.transparent_materials : nullptr;
    const uint64_t* cut = SKIPPING && any_cutout(s) ? s.cutout_materials : nullptr;
    for (uint32_t i = 0; i < s.instance_count; i++) {
        const PyrInstanceData& in = inst[i];
        if (!(in.flags & PYR_INSTANCE_ACTIVE) || !(in.mask & ray_mask)) continue;
        float lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = (in.bounds_min[a] - o[a]) * inv[a];
            hi[a] = (in.bounds_max[a] - o[a]) * inv[a];
        }
        const float t0 = fmaxf(fmaxf(fminf(lo[0], hi[0]), fminf(lo[1], hi[1])), fminf(lo[2], hi[2]));
        const float t1 = fminf(fminf(fmaxf(lo[0], hi[0]), fmaxf(lo[1], hi[1])), fmaxf(lo[2], hi[2]));
        if (!(t0 <= t1)) continue;
        if (!(fmaxf(t0, tmin) < fminf(t1, best))) continue;
        const PyrGeometryData& g = geo[in.geometry];
        if (g.log2_size > levels) continue;  // cannot happen: the stack is sized to the scene's largest edge
        DagView v;
        v.nodes = (const uint32_t*)(uintptr_t)s.nodes + g.node_offset;
        v.leaves = (const uint64_t*)(uintptr_t)s.leaves + g.leaf_offset;
        v.attrs = (g.flags & PYR_GEOMETRY_HAS_ATTRIBUTES) ? (const uint32_t*)(uintptr_t)s.attributes + g.attribute_offset
                                                         : nullptr;
        v.palette = (g.flags & PYR_GEOMETRY_PALETTE) ? (const uint32_t*)(uintptr_t)s.attributes + g.palette_offset
                                                     : nullptr;
        v.root = g.root;
        v.log2_size = g.log2_size;
        v.default_attribute = g.default_attribute;
        v.attr_base = 0;
        float oo[3], dv[3];
        xform_point(in.world_to_object, o, oo);
        xform_vector(in.world_to_object, d, dv);
        DagHitD h;
        // with both sets null the skipping walk returns what the plain one does
        const uint64_t* icut = cut && cutout_here(in) ? cut : nullptr;
        if (walk<SKIPPING>(v, oo, dv, tmin, best, want_attr, stk, &h, skip, icut)) {
            best = h.t;
            result = to_hit(true, h, i);
            if (EXT)
#pragma unroll
                for (int a = 0; a < 3; a++) p_object[a] = oo[a] + dv[a] * h.t;
            if (flags & PYR_TRACE_ANY_HIT) return result;
        }
    }
    return result;
}

// ---- kernels -------------------------------------------------------------------
// Every kernel takes its stack from dynamic LDS: stack_bytes(levels) per block.
extern __shared__ uint32_t s_stk[];

__host__ __device__ constexpr uint32_t stack_bytes(uint32_t levels) { return levels * kTraceBlock * 4u; }

template <bool SKIPPING>
__global__ void __launch_bounds__(kTraceBlock) k_trace_scene(const PyrScene* scene, const PyrRay* rays, PyrHit* hits,
                                                              uint32_t count, uint32_t ray_mask, uint32_t flags,
                                                              uint32_t levels) {
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= count) return;
    const PyrRay r = rays[i];
    hits[i] = trace_scene<SKIPPING>(*scene, r.origin, r.direction, r.tmin, r.tmax, ray_mask, flags, s_stk + threadIdx.x,
                                    levels);
}

// extendedHit and hitNormal (trace.zig:125-157): the voxel from the object
// space hit point moved half a voxel inward on the hit face's axis (unless the
// ray started inside), the hit point o + d t in the rays' space, and the face
// normal through the transpose of world_to_object (vec.zig:43-49), normalised
// as v * (1 / sqrt((x x + y y) + z z)) (vec.zig:13-23). A miss is all zero
// but for the miss hit.
__device__ __forceinline__ void store_extended(const PyrScene& s, const float o[3], const float d[3], const PyrHit& h,
                                               const float po[3], PyrHitEx* out) {
    PyrHitEx r = {};
    r.hit = h;
    if (h.instance != PYR_NO_HIT) {
        const uint32_t face = h.meta & PYR_HIT_FACE_MASK, axis = face >> 1;
        const bool inside = (h.meta & PYR_HIT_INSIDE) != 0;
        float n[3];
#pragma unroll
        for (uint32_t a = 0; a < 3; a++) {
            float p = po[a];
            if (!inside && a == axis) p += (face & 1u) ? 0.5f : -0.5f;
            r.voxel[a] = (int32_t)floorf(p);
            r.position[a] = o[a] + d[a] * h.t;
            n[a] = a == axis ? ((face & 1u) ? -1.0f : 1.0f) : 0.0f;
        }
        const float* m = ((const PyrInstanceData*)(uintptr_t)s.instances)[h.instance].world_to_object;
        const float v0 = m[0] * n[0] + m[4] * n[1] + m[8] * n[2];
        const float v1 = m[1] * n[0] + m[5] * n[1] + m[9] * n[2];
        const float v2 = m[2] * n[0] + m[6] * n[1] + m[10] * n[2];
        const float il = 1.0f / sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
        r.normal[0] = v0 * il;
        r.normal[1] = v1 * il;
        r.normal[2] = v2 * il;
    }
    *out = r;
}

// pyr_trace with PYR_TRACE_EXTENDED (gpu_kernels.zig:79-81): one PyrHitEx per ray
template <bool SKIPPING>
__global__ void __launch_bounds__(kTraceBlock) k_trace_scene_ex(const PyrScene* scene, const PyrRay* rays,
                                                                 PyrHitEx* hits, uint32_t count, uint32_t ray_mask,
                                                                 uint32_t flags, uint32_t levels) {
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= count) return;
    const PyrRay r = rays[i];
    float po[3] = {0.0f, 0.0f, 0.0f};
    const PyrHit h = trace_scene<SKIPPING, true>(*scene, r.origin, r.direction, r.tmin, r.tmax, ray_mask, flags,
                                                  s_stk + threadIdx.x, levels, po);
    store_extended(*scene, r.origin, r.direction, h, po, hits + i);
}

__global__ void __launch_bounds__(kTraceBlock) k_trace_dag(DagView g, const PyrRay* rays, PyrHit* hits,
                                                            uint32_t count, uint32_t flags) {
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= count) return;
    const PyrRay r = rays[i];
    DagHitD h;
    bool f = walk<false>(g, r.origin, r.direction, r.tmin, r.tmax, (flags & PYR_TRACE_NO_ATTRIBUTE) == 0,
                  s_stk + threadIdx.x, &h);
    hits[i] = to_hit(f, h, 0);
}

// cameraRay (render.zig:16-32) for pixel centre + jitter (render.zig:133-141)
__device__ __forceinline__ void camera_ray(const PyrCamera& c, uint32_t x, uint32_t y, float o[3], float d[3],
                                           float* tmin, float* tmax) {
    const float px = (float)x + 0.5f + c.jitter[0];
    const float py = (float)y + 0.5f + c.jitter[1];
    const float nx = 2.0f * px / (float)c.width - 1.0f;
    const float ny = 1.0f - 2.0f * py / (float)c.height;
    const float vx = (nx + c.shift[0]) * c.scale[0];
    const float vy = (ny + c.shift[1]) * c.scale[1];
    const bool ortho = c.projection == PYR_PROJECTION_ORTHOGRAPHIC;
    const float ov[3] = {ortho ? vx : 0.0f, ortho ? vy : 0.0f, 0.0f};
    const float dv[3] = {ortho ? 0.0f : vx, ortho ? 0.0f : vy, -1.0f};
    float dw[3];
    xform_vector(c.view_to_world, dv, dw);
    const float len = sqrtf(dw[0] * dw[0] + dw[1] * dw[1] + dw[2] * dw[2]);
    const float il = 1.0f / len;
    xform_point(c.view_to_world, ov, o);
    d[0] = dw[0] * il;
    d[1] = dw[1] * il;
    d[2] = dw[2] * il;
    *tmin = c.near_plane * len;
    *tmax = c.far_plane > 0 ? c.far_plane * len : PYR_FLT_MAX;
}

// 8x16 pixel tiles per workgroup keep a wave's rays coherent
__global__ void __launch_bounds__(kTraceBlock) k_trace_camera(DagView g, PyrCamera cam, PyrHit* hits,
                                                               uint32_t flags) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 8 + (threadIdx.x >> 4);
    if (x >= cam.width || y >= cam.height) return;
    float o[3], d[3], tmin, tmax;
    camera_ray(cam, x, y, o, d, &tmin, &tmax);
    DagHitD h;
    bool f = walk<false>(g, o, d, tmin, tmax, (flags & PYR_TRACE_NO_ATTRIBUTE) == 0, s_stk + threadIdx.x, &h);
    hits[(uint64_t)y * cam.width + x] = to_hit(f, h, 0);
}

__global__ void __launch_bounds__(kTraceBlock) k_camera_rays(PyrCamera cam, PyrRay* rays) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 8 + (threadIdx.x >> 4);
    if (x >= cam.width || y >= cam.height) return;
    PyrRay r;
    camera_ray(cam, x, y, r.origin, r.direction, &r.tmin, &r.tmax);
    rays[(uint64_t)y * cam.width + x] = r;
}

template <bool SKIPPING>
__global__ void __launch_bounds__(kTraceBlock) k_render_scene(const PyrScene* scene, PyrCamera cam, const float* w2v,
                                                               PyrHit* hits, float* depth, uint32_t ray_mask,
                                                               uint32_t flags, uint32_t levels) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 8 + (threadIdx.x >> 4);
    if (x >= cam.width || y >= cam.height) return;
    float o[3], d[3], tmin, tmax;
    camera_ray(cam, x, y, o, d, &tmin, &tmax);
    PyrHit h = trace_scene<SKIPPING>(*scene, o, d, tmin, tmax, ray_mask, flags, s_stk + threadIdx.x, levels);
    const uint64_t i = (uint64_t)y * cam.width + x;
    if (hits) hits[i] = h;
    if (depth) {
        float dp = PYR_FLT_MAX;
        if (h.instance != PYR_NO_HIT) {
            float p[3], q[3];
#pragma unroll
            for (int a = 0; a < 3; a++) p[a] = o[a] + d[a] * h.t;
            xform_point(w2v, p, q);
            dp = -q[2];
        }
        depth[i] = dp;
    }
}

// shadow ray per primary hit (shade.zig:626-629, 940): from the hit point
// moved 1e-3 voxel along the face normal toward the sun, any hit
__global__ void __launch_bounds__(kTraceBlock) k_trace_shadow(DagView g, const PyrRay* rays, const PyrHit* prim,
                                                               float sx, float sy, float sz, uint32_t* out,
                                                               uint32_t count) {
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= count) return;
    const PyrHit h = prim[i];
    if (h.instance == PYR_NO_HIT) {
        out[i] = 0xFFFFFFFFu;
        return;
    }
    const PyrRay r = rays[i];
    const uint32_t face = h.meta & PYR_HIT_FACE_MASK, axis = face >> 1;
    const float sign = (face & 1u) ? -1.0f : 1.0f;
    float o[3];
    const float d[3] = {sx, sy, sz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float nrm = (uint32_t)a == axis ? sign : 0.0f;
        o[a] = r.origin[a] + r.direction[a] * h.t + nrm * 1e-3f;
    }
    DagHitD sh;
    out[i] = walk<false>(g, o, d, 0.0f, PYR_FLT_MAX, false, s_stk + threadIdx.x, &sh) ? 1u : 0u;
}

DagView view_of(const PyrhipDag& d) {
    DagView g;
    g.nodes = (const uint32_t*)(uintptr_t)d.nodes;
    g.leaves = (const uint64_t*)(uintptr_t)d.leaves;
    g.attrs = (const uint32_t*)(uintptr_t)d.attrs;
    g.palette = d.attrs ? (const uint32_t*)(uintptr_t)d.palette : nullptr;
    g.root = d.root;
    g.log2_size = d.log2_size;
    g.default_attribute = d.default_attribute;
    g.attr_base = 0;
    return g;
}

PyrResult launch_trace_scene(const PyrScene* scene_dev, uint64_t rays, uint64_t hits, uint32_t count,
                             uint32_t ray_mask, uint32_t flags, uint32_t levels, bool skipping, hipStream_t s) {
    if (count == 0) return PYR_OK;
    if (levels < 3 || levels > kMaxLevels) return fail(PYR_ERROR_INVALID_ARGUMENT, "scene levels %u", levels);
    const dim3 grid((count + kTraceBlock - 1) / kTraceBlock);
    if (flags & PYR_TRACE_EXTENDED) {
        if (skipping)
            hipLaunchKernelGGL(k_trace_scene_ex<true>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev,
                               (const PyrRay*)(uintptr_t)rays, (PyrHitEx*)(uintptr_t)hits, count, ray_mask, flags,
                               levels);
        else
            hipLaunchKernelGGL(k_trace_scene_ex<false>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev,
                               (const PyrRay*)(uintptr_t)rays, (PyrHitEx*)(uintptr_t)hits, count, ray_mask, flags,
                               levels);
    } else if (skipping)
        hipLaunchKernelGGL(k_trace_scene<true>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev,
                           (const PyrRay*)(uintptr_t)rays, (PyrHit*)(uintptr_t)hits, count, ray_mask, flags, levels);
    else
        hipLaunchKernelGGL(k_trace_scene<false>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev,
                           (const PyrRay*)(uintptr_t)rays, (PyrHit*)(uintptr_t)hits, count, ray_mask, flags, levels);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "trace launch: %s", hipGetErrorString(e));
}

PyrResult launch_render_scene(const PyrScene* scene_dev, const PyrCamera& cam, const float* w2v_dev, uint64_t hits,
                              uint64_t depth, uint32_t ray_mask, uint32_t flags, uint32_t levels, bool skipping,
                              hipStream_t s) {
    if (levels < 3 || levels > kMaxLevels) return fail(PYR_ERROR_INVALID_ARGUMENT, "scene levels %u", levels);
    dim3 grid((cam.width + 15) / 16, (cam.height + 7) / 8);
    if (skipping)
        hipLaunchKernelGGL(k_render_scene<true>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev, cam,
                           w2v_dev, (PyrHit*)(uintptr_t)hits, (float*)(uintptr_t)depth, ray_mask, flags, levels);
    else
        hipLaunchKernelGGL(k_render_scene<false>, grid, dim3(kTraceBlock), stack_bytes(levels), s, scene_dev, cam,
                           w2v_dev, (PyrHit*)(uintptr_t)hits, (float*)(uintptr_t)depth, ray_mask, flags, levels);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "render launch: %s", hipGetErrorString(e));
}

}  // namespace pyr

using namespace pyr;

extern "C" {

PYR_API PyrResult pyrhip_trace_dag(const PyrhipDag* dag, uint64_t rays, uint64_t hits, uint32_t count,
                                   uint32_t flags, void* stream) {
    diag_clear();
    if (!dag || (count && (!rays || !hits))) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->log2_size < 3 || dag->log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size");
    if (count == 0) return PYR_OK;
    hipLaunchKernelGGL(k_trace_dag, dim3((count + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock),
                       stack_bytes(dag->log2_size), (hipStream_t)stream, view_of(*dag), (const PyrRay*)(uintptr_t)rays, (PyrHit*)(uintptr_t)hits,
                       count, flags);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "trace launch: %s", hipGetErrorString(e));
}

PYR_API PyrResult pyrhip_trace_camera(const PyrhipDag* dag, const PyrCamera* camera, uint64_t hits, uint32_t flags,
                                      void* stream) {
    diag_clear();
    if (!dag || !camera || !hits) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->log2_size < 3 || dag->log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size");
    if (camera->width == 0 || camera->height == 0) return PYR_OK;
    dim3 grid((camera->width + 15) / 16, (camera->height + 7) / 8);
    hipLaunchKernelGGL(k_trace_camera, grid, dim3(kTraceBlock), stack_bytes(dag->log2_size), (hipStream_t)stream,
                       view_of(*dag), *camera,
                       (PyrHit*)(uintptr_t)hits, flags);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "trace launch: %s", hipGetErrorString(e));
}

PYR_API PyrResult pyrhip_camera_rays(const PyrCamera* camera, uint64_t rays, void* stream) {
    diag_clear();
    if (!camera || !rays) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (camera->width == 0 || camera->height == 0) return PYR_OK;
    dim3 grid((camera->width + 15) / 16, (camera->height + 7) / 8);
    hipLaunchKernelGGL(k_camera_rays, grid, dim3(kTraceBlock), 0, (hipStream_t)stream, *camera,
                       (PyrRay*)(uintptr_t)rays);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "camera rays: %s", hipGetErrorString(e));
}

PYR_API PyrResult pyrhip_trace_shadow(const PyrhipDag* dag, uint64_t rays, uint64_t hits, const float sun[3],
                                      uint64_t out, uint32_t count, void* stream) {
    diag_clear();
    if (!dag || !sun || (count && (!rays || !hits || !out))) return fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    if (dag->log2_size < 3 || dag->log2_size > 20) return fail(PYR_ERROR_INVALID_ARGUMENT, "log2_size");
    if (count == 0) return PYR_OK;
    hipLaunchKernelGGL(k_trace_shadow, dim3((count + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock),
                       stack_bytes(dag->log2_size), (hipStream_t)stream, view_of(*dag), (const PyrRay*)(uintptr_t)rays,
                       (const PyrHit*)(uintptr_t)hits, sun[0], sun[1], sun[2], (uint32_t*)(uintptr_t)out, count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PYR_OK : fail(PYR_ERROR_CUDA, "trace launch: %s", hipGetErrorString(e));
}

}  // extern "C"
