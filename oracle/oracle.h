/*
 * oracle.h - CPU restatement of Pyrit's sparse-voxel-DAG build and software
 * traversal. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it as the checker. Nothing in the product
 * (pyrit_amd/, libpyrit) may link, load or call it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root). The reference is Zig 0.16 and cannot be
 * compiled here, so this is written from the source text.
 *
 * Pinning (what checks each part against something other than itself):
 *   morton / decode     : known answers dag_builder.zig:420-426 + roundtrip
 *   or_build_points/..  : empty/single-voxel KATs (dag_builder.zig:428-442),
 *                         dense == points bit-identical (:444-498), period-8
 *                         dedup bound (:500-520), per-voxel lookup against the
 *                         dense input (brute force), minimality (no two equal
 *                         node tuples) checked in tests
 *   or_gb_build*        : canon(or_gb_build(P)) == or_build_points(P) and the
 *                         lookup/attribute equalities of cpu_test.zig:758-850,
 *                         chunk batch == per-chunk single (cpu_test.zig:1001-1060)
 *   or_canon            : canon(or_build_points(P)) == or_build_points(P)
 *   or_walk_skipping    : skipping a material == the same walk on the DAG
 *                         built without that material's voxels (hits bit for
 *                         bit) and == the f64 DDA with those voxels empty;
 *                         empty sets == or_walk; cutouts: a cutout instance
 *                         scaled past the voxel-edge limit == no cutout, holes
 *                         on about 30 % of entry points (77/256); the hole
 *                         hash constants themselves are parity unpinned (the
 *                         reference has no test of them)
 *   or_walk / scene     : axis-ray KATs (cpu_test.zig:15-40), f64 DDA on a dense
 *                         grid (tests/common.zig:104-153, cpu_test.zig:42-103)
 *   or_camera_*         : closed forms (centre pixel looks at target; unit dirs)
 *   or_trace_rays with  : three transformed, masked instances against the f64
 *   transforms            DDA in each instance's object space (cpu_test.zig:
 *                         105-171): nearest instance and voxel attribute, hit
 *                         point on the voxel face within 2e-3
 *   or_inverse          : known answer and singular case (xform.zig:60-68),
 *                         equal to the f64 LU inverse on random transforms
 *   or_transformed_box  : the box of the 8 transformed corners widened by the
 *                         stated margin (brute force)
 *   or_render_primary   : depth = t / |(vx, vy, -1)| (closed form of the
 *                         pinhole view-space depth), hits = plain trace
 *   or_range_*          : the allocation sequence of range_alloc.zig:61-78
 *   or_palette_pack     : decoding the packed words reproduces every attribute
 *                         (brute force), the 16-value boundary and the word
 *                         counts of context.zig:1352-1357
 *   or_rt_prims         : known answer rt_prims.zig:139-160; brute force from
 *                         the voxel set (cells = occupied 2^k cubes in Morton
 *                         order, box = min/max of their voxels, attr_base =
 *                         voxels before the cell in Morton order, node = a
 *                         subtree whose lookups give exactly the cell's voxels)
 *   or_trace_rays_ex    : the voxel equals the walk's own hit voxel (the
 *                         start voxel for inside hits) but at float grazing
 *                         edges; hits equal or_trace_rays; the normal of a
 *                         similarity instance equals object_to_world times
 *                         the face normal, normalised in f64 (within 1e-6),
 *                         and faces the ray (dot(normal, d) < 0) for hits
 *                         from outside; misses are zero but for the hit
 *   or_trace_touch      : word counts derived by hand on a one-voxel DAG (hit,
 *                         miss through the same leaf, ray outside the bounds),
 *                         per-ray distinctness, hit count equal to the tracer's
 *                         (measurement instrument for the traversal's
 *                         algorithmic bytes, not a checker)
 */
#ifndef PYRIT_ORACLE_H
#define PYRIT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrVoxel { int32_t x, y, z; uint32_t attribute; } OrVoxel;

/* CPU builder result (dag_builder.zig:18-31) */
typedef struct OrDag {
    uint32_t log2_size, root;
    uint64_t voxel_count;
    uint32_t* nodes;  uint64_t node_words;
    uint64_t* leaves; uint64_t leaf_count;
    uint32_t* attrs;  uint64_t attr_count; /* NULL when not stored */
} OrDag;

/* serial GPU-build result in the GPU-native layout (gpu_build.zig:25-47) */
typedef struct OrBuilt {
    uint32_t log2_size;
    uint32_t root;
    uint32_t* nodes;  uint32_t node_words;
    uint64_t* leaves; uint32_t leaf_count;
    uint64_t* keys;   uint32_t* attrs; uint32_t voxel_count;
} OrBuilt;

uint64_t or_morton(uint32_t x, uint32_t y, uint32_t z);
void     or_morton_decode(uint64_t m, uint32_t out[3]);

/* --- CPU builder (dag_builder.zig) ---------------------------------------- */
int      or_build_points(uint32_t log2_size, const OrVoxel* pts, size_t count, int store_attrs, OrDag* out);
int      or_build_dense(uint32_t log2_size, const uint32_t* voxels, int store_attrs, OrDag* out);
void     or_dag_free(OrDag* d);
/* attribute at (x,y,z), 0 when empty (Dag.lookup, dag_builder.zig:35-61) */
uint32_t or_lookup(const uint32_t* nodes, const uint64_t* leaves, const uint32_t* attrs, uint32_t root,
                   uint32_t log2_size, uint32_t x, uint32_t y, uint32_t z);

/* --- serial GPU build (gbuild.zig ops in gpu_build.zig:135-360 order) ------ */
int  or_gb_build(uint32_t log2_size, const uint64_t* old_keys, const uint32_t* old_vals, uint32_t n_old,
                 const OrVoxel* edits, uint32_t n_new, OrBuilt* out);
/* Chunk batch (gpu_build.zig:98-122). chunk_log2 may be NULL (all chunks
 * log2_size); otherwise chunk c has edge 2^chunk_log2[c] <= 2^log2_size (this
 * repository's mixed-size extension). capacity 0: packed layout, chunk c's
 * voxels at offsets[c] (also an extension). roots/first_voxel get 0xFFFFFFFF
 * for empty chunks. */
int  or_gb_build_chunks(uint32_t log2_size, const uint8_t* chunk_log2, uint32_t count, uint32_t capacity,
                        const OrVoxel* voxels, const uint32_t* offsets, OrBuilt* out, uint32_t* roots,
                        uint32_t* first_voxel);
void or_built_free(OrBuilt* b);

/* Canonical form: DFS from root, post-order interning of remapped tuples in one
 * cross-level table, leaves merged by mask and renumbered by first visit
 * (reproduces dag_builder; a DAG with repeated leaves or tuples, as path edits
 * leave behind, maps to the minimal DAG of the same voxels). */
int  or_canon(const uint32_t* nodes, const uint64_t* leaves, uint32_t root, uint32_t log2_size, OrDag* out);

/* --- traversal (device/dag.zig, device/trace.zig, device/render.zig) ------- */
typedef struct OrDagHit { float t; uint32_t face; int32_t voxel[3]; uint32_t attribute; uint32_t inside; } OrDagHit;
typedef struct OrDagView {
    const uint32_t* nodes; const uint64_t* leaves; const uint32_t* attrs;
    uint32_t root, log2_size, default_attribute, attr_base;
} OrDagView;
/* 1 = hit (dag.zig:160-351, exit_mode = false, no skip/cut) */
int  or_walk(const OrDagView* g, const float o[3], const float d[3], float tmin, float tmax, int want_attr,
             OrDagHit* hit, uint32_t* steps_out);
/* traceSkipping (dag.zig:113-115, 133-157, 236-289): voxels whose material
 * (attribute & 0xFF) is in `skip` are passed; voxels whose material is in
 * `cut` are passed where the entry point falls in a hole of the material's
 * fixed pattern (about 30 % of each face), or when the ray starts inside
 * them. Either set may be NULL. With a set given, hits carry the attribute
 * even when want_attr is 0. */
int  or_walk_skipping(const OrDagView* g, const float o[3], const float d[3], float tmin, float tmax, int want_attr,
                      const uint64_t skip[4], const uint64_t cut[4], const int32_t base[3], OrDagHit* hit);

/* Layouts identical to PyrRay / PyrHit / PyrGeometryData / PyrInstanceData */
typedef struct OrRay { float origin[3]; float tmin; float direction[3]; float tmax; } OrRay;
typedef struct OrHit { float t; uint32_t instance, attribute, meta; } OrHit;
typedef struct OrGeometry {
    uint32_t node_offset, leaf_offset, attribute_offset, root, log2_size, flags, default_attribute, palette_offset;
} OrGeometry;
typedef struct OrInstance {
    float object_to_world[12]; float world_to_object[12];
    float bounds_min[3]; uint32_t geometry; float bounds_max[3]; uint32_t mask;
    uint32_t user, history, flags, reserved;
} OrInstance;
typedef struct OrScene {
    const uint32_t* nodes; const uint64_t* leaves; const uint32_t* attrs;
    const OrGeometry* geometries; const OrInstance* instances; uint32_t instance_count;
    uint64_t transparent[4], cutout[4]; /* material sets (PyrScene transparent/cutout_materials) */
} OrScene;
/* pyr_trace semantics over a ray array (trace.zig:72-123); flags 0x8 =
 * PYR_TRACE_SKIP_TRANSPARENT; cutout materials apply on instances whose voxel
 * edge is at most about 1.4 world units */
void or_trace_rays(const OrScene* s, const OrRay* rays, OrHit* hits, uint64_t count, uint32_t ray_mask,
                   uint32_t flags, uint32_t threads);

/* PYR_TRACE_EXTENDED hits (trace.zig:125-157; the layout of PyrHitEx) */
typedef struct OrHitEx {
    OrHit hit; int32_t voxel[3]; uint32_t reserved0; float position[3]; uint32_t reserved1;
    float normal[3]; uint32_t reserved2;
} OrHitEx;
void or_trace_rays_ex(const OrScene* s, const OrRay* rays, OrHitEx* hits, uint64_t count, uint32_t ray_mask,
                      uint32_t flags, uint32_t threads);

/* Instrumented trace (one thread): sums over the rays of the distinct node
 * words, leaves and attributes each ray's traversal reads, and the hit count:
 * out = {node words, leaves, attributes, hits}. The per-ray distinct bytes
 * of the traversal's algorithmic-byte model. */
void or_trace_touch(const OrScene* s, const OrRay* rays, uint64_t count, uint32_t ray_mask, uint32_t flags,
                    uint64_t out[4]);

/* Camera (types.zig:70-86) and helpers (root.zig:449-481, render.zig:16-32) */
typedef struct OrCamera {
    float view_to_world[12]; uint32_t projection, width, height;
    float scale[2], shift[2], jitter[2], near_plane, far_plane;
} OrCamera;
void or_camera_look_at(OrCamera* c, const float eye[3], const float target[3], const float up[3]);
void or_camera_perspective(OrCamera* c, float fov_y, uint32_t width, uint32_t height, float near_plane);
/* primary ray of pixel (x,y) (render.zig:133-141), one ray per pixel row-major */
void or_camera_rays(const OrCamera* c, OrRay* rays);
int  or_inverse(const float m[12], float out[12]);                       /* xform.zig:8-34 */
void or_transformed_box(const float m[12], float size, float lo[3], float hi[3]); /* xform.zig:37-58 */
/* primary hits + linear depth (render.zig:143-166), ray_mask = all */
void or_render_primary(const OrScene* s, const OrCamera* c, OrHit* hits, float* depth, uint32_t threads);

/* f64 DDA on a dense grid (tests/common.zig:104-153). Returns 1 on hit. */
int or_ref_dda(const uint32_t* grid, uint32_t n, const double o[3], const double d[3], double tmin, double tmax,
               double* t, int32_t v[3], int32_t* face);

/* range allocator (range_alloc.zig:10-59) */
typedef struct OrRange { uint64_t offset, size; } OrRange;
typedef struct OrRangeAlloc { uint64_t capacity, used; OrRange* free; uint32_t n, cap; } OrRangeAlloc;
void     or_range_init(OrRangeAlloc* r, uint64_t capacity);
int64_t  or_range_alloc(OrRangeAlloc* r, uint64_t size); /* -1 = full */
void     or_range_release(OrRangeAlloc* r, uint64_t offset, uint64_t size);
void     or_range_free(OrRangeAlloc* r);

/* palette packing of a batch's attributes (palette.zig:1-152, layout of
 * context.zig:1311-1370): chunk j owns ranks [first[j], first[j+1]);
 * count[j] = distinct values, or 0xFFFFFFFF above 16 (or when 0xFFFFFFFF
 * occurs); word_offset[j] its start (word_offset[k] = total). A palette chunk
 * stores 16 entries, then the index of rank r in bits 4 (r & 7) of word
 * r >> 3; the entries are the distinct values in ascending order padded with
 * the first (the reference keeps hash-slot order, which concurrent inserts
 * leave unspecified; any order decodes to the same attributes). out == NULL:
 * layout only. */
void or_palette_pack(const uint32_t* attrs, const uint32_t* first, uint32_t k, uint32_t* count,
                     uint64_t* word_offset, uint32_t* out);

/* RT primitives (rt_prims.zig:47-137 extract; the RtPrim of types.zig:
 * 518-525): every non-empty subtree of edge 2^k, k = clamp(requested, 3,
 * log2_size), in DFS order: {node word, DFS rank of its first voxel, cell
 * x | y << 21 | z << 42 in units of 2^k} and its voxel box (max exclusive)
 * widened by 1/256, aabbs[6 i ..] = min xyz, max xyz. Writes the first
 * `capacity` primitives (prims/aabbs may be NULL), returns the total count, -1
 * when memory runs out; *out_k = k. */
typedef struct OrRtPrim { uint32_t node, attr_base; uint64_t cell; } OrRtPrim;
int64_t or_rt_prims(const uint32_t* nodes, const uint64_t* leaves, uint32_t root, uint32_t log2_size,
                    uint32_t requested_log2, OrRtPrim* prims, float* aabbs, uint64_t capacity, uint32_t* out_k);

#ifdef __cplusplus
}
#endif
#endif
