/*
 * pyrit.h - C ABI of the MI355X-native Pyrit hot path (libpyrit.so).
 *
 * This is the subset of the reference's public header (include/pyrit.h of
 * SimonVitzethum/Pyrit) that the GPU sparse-voxel-DAG build and the software
 * DAG traversal sit behind. Struct layouts, constants, names and argument
 * meanings are kept identical so a caller of the reference links against this
 * library unchanged; each block cites the reference lines it mirrors. The
 * implementation is HIP on gfx950; "cuda" in two legacy names below means the
 * HIP stream of the context.
 *
 * Context functions are not thread-safe (reference pyrit.h:27-28). All GPU
 * work is asynchronous on the context stream (pyr_cuda_stream).
 */
#ifndef PYRIT_H
#define PYRIT_H

#include <stddef.h>
#if defined(__cplusplus)
#include <cstdint>
#else
#include <stdint.h>
#endif

#define PYR_FLT_MAX 3.402823466e+38f

/* ---- rays and hits (reference pyrit.h:46-91) ----------------------------- */

typedef struct PyrRay {
    float origin[3];
    float tmin;
    float direction[3]; /* need not be normalised; t is measured in its units */
    float tmax;
} PyrRay;

#define PYR_NO_HIT 0xFFFFFFFFu

/* face of the hit voxel in object space */
#define PYR_FACE_POS_X 0u
#define PYR_FACE_NEG_X 1u
#define PYR_FACE_POS_Y 2u
#define PYR_FACE_NEG_Y 3u
#define PYR_FACE_POS_Z 4u
#define PYR_FACE_NEG_Z 5u

#define PYR_HIT_FACE_MASK 0x7u
#define PYR_HIT_NEW 0x8u
#define PYR_HIT_NO_HISTORY 0x10u
#define PYR_HIT_INSIDE 0x20u /* ray starts inside a filled voxel: t = tmin */

typedef struct PyrHit {
    float t;            /* PYR_FLT_MAX on a miss */
    uint32_t instance;  /* instance index or PYR_NO_HIT */
    uint32_t attribute; /* voxel attribute */
    uint32_t meta;      /* face (PYR_HIT_FACE_MASK) | PYR_HIT_* */
} PyrHit;

#define PYR_TRACE_ANY_HIT 0x1u
#define PYR_TRACE_NO_ATTRIBUTE 0x2u
#define PYR_TRACE_EXTENDED 0x4u /* pyr_trace writes PyrHitEx (voxel, position, normal), e.g. for picking */
#define PYR_TRACE_SKIP_TRANSPARENT 0x8u

/* extended hit (reference pyrit.h:80-88) */
typedef struct PyrHitEx {
    PyrHit hit;
    int32_t voxel[3];   /* voxel coordinate in the geometry's object space */
    uint32_t reserved0;
    float position[3];  /* hit point in world space (relative to the render origin) */
    uint32_t reserved1;
    float normal[3];    /* world normal */
    uint32_t reserved2;
} PyrHitEx;

/* ---- camera (reference pyrit.h:96-113) ----------------------------------- */

#define PYR_PROJECTION_PERSPECTIVE 0u
#define PYR_PROJECTION_ORTHOGRAPHIC 1u

typedef struct PyrCamera {
    float view_to_world[12]; /* 3x4 row-major, column 3 = position; looks along -Z, +Y up */
    uint32_t projection;
    uint32_t width, height;
    float scale[2];  /* perspective: tan(fov_x/2), tan(fov_y/2) */
    float shift[2];  /* lens shift in NDC units */
    float jitter[2]; /* sub-pixel offset in pixels */
    float near_plane;
    float far_plane; /* 0 = infinite */
} PyrCamera;

/* ---- scene on the device (reference pyrit.h:119-160) --------------------- */

#define PYR_GEOMETRY_HAS_ATTRIBUTES 0x1u
#define PYR_GEOMETRY_PALETTE 0x2u

typedef struct PyrGeometryData {
    uint32_t node_offset, leaf_offset, attribute_offset, root;
    uint32_t log2_size, flags, default_attribute;
    uint32_t palette_offset;
} PyrGeometryData;

#define PYR_INSTANCE_ACTIVE 0x1u

typedef struct PyrInstanceData {
    float object_to_world[12];
    float world_to_object[12];
    float bounds_min[3];
    uint32_t geometry;
    float bounds_max[3];
    uint32_t mask;
    uint32_t user;
    uint32_t history;
    uint32_t flags;
    uint32_t reserved;
} PyrInstanceData;

typedef struct PyrScene {
    uint64_t nodes, leaves, attributes, geometries; /* device pointers */
    uint64_t instances, instances_prev;
    uint32_t instance_count, geometry_count;
    uint64_t frame;
    double time, time_prev;
    double origin[3], origin_prev[3];
    uint64_t materials;
    uint64_t lighting;
    uint64_t transparent_materials[4];
    uint64_t cutout_materials[4];
    uint64_t textures;
    uint32_t texture_count;
    uint32_t reserved_tex;
} PyrScene;

/* ---- results (reference pyrit.h:320-353) --------------------------------- */

#define PYR_VERSION_MAJOR 0
#define PYR_VERSION_MINOR 2
#define PYR_VERSION ((uint32_t)((PYR_VERSION_MAJOR << 16) | PYR_VERSION_MINOR))

#define PYR_API __attribute__((visibility("default")))

#define PYR_DEFINE_HANDLE(name) typedef struct name##_T* name

typedef struct PyrContext PyrContext;
PYR_DEFINE_HANDLE(PyrGeometry);
PYR_DEFINE_HANDLE(PyrInstance);

typedef int32_t PyrResult;
#define PYR_OK 0
#define PYR_ERROR_INVALID_ARGUMENT -1
#define PYR_ERROR_INVALID_HANDLE -2
#define PYR_ERROR_OUT_OF_MEMORY -3
#define PYR_ERROR_CAPACITY -4
#define PYR_ERROR_IN_USE -5
#define PYR_ERROR_CUDA -6 /* a HIP runtime call failed */
#define PYR_ERROR_COMPILE -7
#define PYR_ERROR_NOT_FOUND -8
#define PYR_ERROR_VERSION -9

#ifdef __cplusplus
extern "C" {
#endif

PYR_API const char* pyr_result_string(PyrResult result);
/* message of the last error on this thread (cleared at each API entry) */
PYR_API const char* pyr_error_message(void);

/* ---- context (reference pyrit.h:355-407) --------------------------------- */

#define PYR_CREATE_DEBUG 0x1u
#define PYR_CREATE_NO_RT 0x2u
#define PYR_CREATE_FORCE_RT 0x4u

typedef void (*PyrLogFn)(void* user, int32_t level, const char* message);

typedef struct PyrCreateInfo {
    uint32_t struct_size; /* sizeof(PyrCreateInfo) */
    uint32_t version;     /* PYR_VERSION */
    int32_t device;       /* HIP device index */
    uint32_t flags;       /* PYR_CREATE_* */
    void* cuda_context;   /* ignored (HIP has no explicit contexts) */
    void* cuda_stream;    /* existing hipStream_t or NULL (own stream) */
    uint32_t max_instances;  /* 0 = 65536 */
    uint32_t max_geometries; /* 0 = 4096 */
    uint32_t max_views;      /* 0 = 16 */
    uint32_t max_textures;   /* 0 = 256 */
    uint32_t rt_leaf_log2;   /* unused on this backend (no RT cores) */
    uint64_t node_pool_bytes;      /* 0 = 256 MiB */
    uint64_t leaf_pool_bytes;      /* 0 = 256 MiB */
    uint64_t attribute_pool_bytes; /* 0 = 256 MiB */
    uint64_t staging_bytes;        /* unused */
    PyrLogFn log;
    void* log_user;
} PyrCreateInfo;

PYR_API PyrResult pyr_create(const PyrCreateInfo* info, PyrContext** out);
PYR_API void pyr_destroy(PyrContext* ctx); /* waits for the GPU */
PYR_API void* pyr_cuda_stream(PyrContext* ctx); /* the context's hipStream_t */
PYR_API uint32_t pyr_features(PyrContext* ctx); /* 0: no RT cores */
PYR_API PyrResult pyr_synchronize(PyrContext* ctx);

/* pool use in bytes (the path-edit arena is not part of the capacity), live
 * geometries and instances, committed frames (reference pyrit.h:712-722) */
typedef struct PyrStats {
    uint64_t node_pool_used, node_pool_capacity;
    uint64_t leaf_pool_used, leaf_pool_capacity;
    uint64_t attribute_pool_used, attribute_pool_capacity;
    uint32_t geometries, instances;
    uint64_t frame;
    uint32_t features;
    uint32_t reserved;
} PyrStats;

PYR_API PyrResult pyr_get_stats(PyrContext* ctx, PyrStats* out);

/* ---- voxels (reference pyrit.h:413-416) ---------------------------------- */

typedef struct PyrVoxel {
    int32_t x, y, z;
    uint32_t attribute; /* != 0; 0 deletes in edits */
} PyrVoxel;

/* ---- host DAGs (reference pyrit.h:401-444) --------------------------------
 *
 * Host-side DAGs in the DAG format (nodes, leaves, attributes in DFS rank
 * order). This implementation builds them on the GPU (the same build as
 * pyr_geometry_build) and downloads the arrays, so their node order is the
 * GPU build's; the reference's CPU builder orders nodes differently, and both
 * describe the same voxels, traverse identically and are equal after
 * canonicalisation. pyr_dag_* functions are thread-safe and need a HIP device. */

#define PYR_DAG_NO_ATTRIBUTES 0x1u /* geometry only; hits report default_attribute = 1 */

typedef struct PyrDag PyrDag;

/* attribute of voxel (x, y, z), 0 = empty */
typedef uint32_t (*PyrVoxelFn)(void* user, int32_t x, int32_t y, int32_t z);
/* optional: is the cube [x, x+size)^3 certainly empty? non-zero skips it */
typedef int32_t (*PyrRegionEmptyFn)(void* user, int32_t x, int32_t y, int32_t z, uint32_t size);

typedef struct PyrDagInfo {
    uint32_t log2_size;
    uint32_t root;
    uint64_t voxel_count;
    const uint32_t* nodes;
    uint64_t node_words;
    const uint64_t* leaves;
    uint64_t leaf_count;
    const uint32_t* attributes; /* NULL with PYR_DAG_NO_ATTRIBUTES */
    uint64_t attribute_count;
} PyrDagInfo;

/* voxels[x + N * (y + N * z)], N = 2^log2_size */
PYR_API PyrResult pyr_dag_build_dense(uint32_t log2_size, const uint32_t* voxels, uint32_t flags, PyrDag** out);
PYR_API PyrResult pyr_dag_build_points(uint32_t log2_size, const PyrVoxel* voxels, size_t count, uint32_t flags,
                                       PyrDag** out);
PYR_API PyrResult pyr_dag_build_fn(uint32_t log2_size, PyrVoxelFn voxel, PyrRegionEmptyFn region_empty, void* user,
                                   uint32_t flags, PyrDag** out);
PYR_API void pyr_dag_get_info(const PyrDag* dag, PyrDagInfo* out);
PYR_API void pyr_dag_destroy(PyrDag* dag);
/* the reference's binary DAG format ("PYRD", src/io.zig:86-160); buffer == NULL: only *size */
PYR_API PyrResult pyr_dag_save(const PyrDag* dag, void* buffer, size_t* size);
PYR_API PyrResult pyr_dag_load(const void* data, size_t size, PyrDag** out);

/* ---- geometry on the GPU (reference pyrit.h:454-480) --------------------- */

/* uploads a host DAG (the DAG may be destroyed right after) */
PYR_API PyrResult pyr_geometry_create(PyrContext* ctx, const PyrDag* dag, PyrGeometry* out);
/* coarser LOD of an editable geometry built on the GPU: edge 2^(log2_size -
 * shift), one voxel per coarse cell (the source voxel with the largest key);
 * flags: PYR_BUILD_EDITABLE */
PYR_API PyrResult pyr_geometry_downsample(PyrContext* ctx, PyrGeometry geometry, uint32_t shift, uint32_t flags,
                                          PyrGeometry* out);
/* copies a geometry from the GPU into a host DAG (e.g. for pyr_dag_save) */
PYR_API PyrResult pyr_geometry_download(PyrContext* ctx, PyrGeometry geometry, PyrDag** out);

#define PYR_BUILD_HOST_INPUT 0x1u /* voxels is a host pointer */
#define PYR_BUILD_EDITABLE 0x2u   /* keep the sorted voxels: pyr_geometry_edit allowed */

/* DAG from PyrVoxel[count] in any order (duplicates: last wins, attribute 0 =
 * empty) built on the GPU; voxels is a device pointer unless HOST_INPUT. */
PYR_API PyrResult pyr_geometry_build(PyrContext* ctx, uint32_t log2_size, const void* voxels, uint32_t count,
                                     uint32_t flags, PyrGeometry* out);
/* set or remove voxels of an editable geometry (rebuild on the GPU) */
PYR_API PyrResult pyr_geometry_edit(PyrContext* ctx, PyrGeometry geometry, const void* voxels, uint32_t count,
                                    uint32_t flags);
PYR_API PyrResult pyr_geometry_destroy(PyrContext* ctx, PyrGeometry geometry); /* IN_USE while instanced */
PYR_API uint32_t pyr_geometry_index(PyrGeometry geometry);

/* ---- instances (reference pyrit.h:482-502) ------------------------------- */

PYR_API PyrResult pyr_instance_create(PyrContext* ctx, PyrGeometry geometry, PyrInstance* out);
PYR_API PyrResult pyr_instance_destroy(PyrContext* ctx, PyrInstance instance);
PYR_API PyrResult pyr_instance_set_transform(PyrContext* ctx, PyrInstance instance, const float object_to_world[12]);
PYR_API PyrResult pyr_instance_set_mask(PyrContext* ctx, PyrInstance instance, uint32_t mask); /* default 0xFF */
/* another geometry for the instance (bounds from its size, a new history);
 * like every instance change it reaches the device at the next commit */
PYR_API PyrResult pyr_instance_set_geometry(PyrContext* ctx, PyrInstance instance, PyrGeometry geometry);
PYR_API PyrResult pyr_instance_set_user(PyrContext* ctx, PyrInstance instance, uint32_t user); /* InstanceData.user */
/* a new InstanceData.history (a teleport: the next frame has no past) */
PYR_API PyrResult pyr_instance_reset_history(PyrContext* ctx, PyrInstance instance);
PYR_API uint32_t pyr_instance_index(PyrInstance instance); /* = PyrHit.instance */

/* ---- frames (reference pyrit.h:504-515) ---------------------------------- */

typedef struct PyrFrameInfo {
    double time;
    double origin[3];
} PyrFrameInfo;

PYR_API PyrResult pyr_commit(PyrContext* ctx, const PyrFrameInfo* info);
PYR_API uint64_t pyr_scene_device(PyrContext* ctx); /* device pointer to PyrScene */

/* ---- rendering and tracing (reference pyrit.h:525-550) ------------------- */

PYR_DEFINE_HANDLE(PyrView);

typedef struct PyrTargets {
    uint64_t hits;     /* PyrHit[width * height] or 0 */
    uint64_t depth;    /* float[width * height]: linear depth along the view axis, PYR_FLT_MAX on a miss */
    uint64_t motion;   /* not produced by this backend (must be 0) */
    uint64_t color;    /* not produced (must be 0) */
    uint64_t normal;   /* not produced (must be 0) */
    uint64_t albedo;   /* not produced (must be 0) */
    uint64_t material; /* not produced (must be 0) */
    uint32_t ray_mask; /* 0 = all instances */
    uint32_t flags;    /* PYR_TRACE_NO_ATTRIBUTE */
    uint32_t transparent_mask;
    uint32_t secondary_mask;
    uint32_t coverage;
    float detail_scale;
} PyrTargets;

PYR_API PyrResult pyr_view_create(PyrContext* ctx, PyrView* out);
PYR_API PyrResult pyr_view_destroy(PyrContext* ctx, PyrView view);
/* views keep no frame history here (no temporal passes): checks the handle */
PYR_API PyrResult pyr_view_reset_history(PyrContext* ctx, PyrView view);
/* primary hits and depth of every pixel */
PYR_API PyrResult pyr_render(PyrContext* ctx, PyrView view, const PyrCamera* camera, const PyrTargets* targets);

/* arbitrary rays: rays and hits are device pointers (PyrRay[count], PyrHit[count],
 * PyrHitEx[count] with PYR_TRACE_EXTENDED); flags PYR_TRACE_ANY_HIT |
 * PYR_TRACE_NO_ATTRIBUTE | PYR_TRACE_SKIP_TRANSPARENT | PYR_TRACE_EXTENDED */
PYR_API PyrResult pyr_trace(PyrContext* ctx, uint64_t rays, uint64_t hits, uint32_t count, uint32_t ray_mask,
                            uint32_t flags);

/* ---- materials (reference pyrit.h:162-228, 556-573) -----------------------
 * A voxel attribute carries its material index in bits 0..7 and an sRGB
 * colour 0xRRGGBB in bits 8..31. This backend shades nothing; of a material
 * it uses the flags the traversal reads: with PYR_TRACE_SKIP_TRANSPARENT rays
 * pass voxels of PYR_MATERIAL_TRANSPARENT materials, and every ray passes the
 * holes of PYR_MATERIAL_CUTOUT materials (a fixed pattern, about 30 % of each
 * voxel face) on instances whose voxel edge is at most about 1.4 world
 * units. Both sets take effect at the next pyr_commit. The whole table is
 * kept on the device for the caller's own kernels (PyrScene.materials). */
#define PYR_MAX_MATERIALS 256u
#define PYR_MATERIAL_VOXEL_COLOR 0x1u
#define PYR_MATERIAL_REFRACT 0x2u
#define PYR_MATERIAL_TRANSPARENT 0x4u
#define PYR_MATERIAL_WAVES 0x8u
#define PYR_MATERIAL_CUTOUT 0x10u
#define PYR_MAX_TRANSPARENT_LAYERS 4u
#define PYR_VOXEL(material, r, g, b) \
    ((((uint32_t)(r) & 0xFFu) << 24) | (((uint32_t)(g) & 0xFFu) << 16) | (((uint32_t)(b) & 0xFFu) << 8) | \
     ((uint32_t)(material) & 0xFFu))

typedef struct PyrMaterial {
    float base_color[3];
    float roughness;
    float emission[3];
    float metallic;
    uint32_t flags; /* PYR_MATERIAL_* */
    float opacity;
    float ior;
    float density;
    float wave_height, wave_length, wave_speed;
    float clearcoat;
    float clearcoat_roughness;
    float subsurface;
    float subsurface_color[3];
    uint32_t texture;
    uint32_t normal_texture;
    float texture_scale;
    float normal_strength;
    float normal_scale;
    uint32_t side_texture;
    float side_color[3];
    float variation[3];
    float variation_scale[2];
    float variation_warm;
} PyrMaterial;

PYR_API void pyr_material_default(PyrMaterial* material);
PYR_API PyrResult pyr_material_set(PyrContext* ctx, uint32_t index, const PyrMaterial* material);
PYR_API uint32_t pyr_voxel_attribute(uint32_t material, uint32_t r, uint32_t g, uint32_t b);

/* ---- camera helpers (reference pyrit.h:917-921) -------------------------- */

PYR_API void pyr_camera_look_at(PyrCamera* camera, const float eye[3], const float target[3], const float up[3]);
PYR_API void pyr_camera_perspective(PyrCamera* camera, float fov_y_radians, uint32_t width, uint32_t height,
                                    float near_plane);
/* parallel rays; the view spans +-half_height world units vertically */
PYR_API void pyr_camera_orthographic(PyrCamera* camera, float half_height, uint32_t width, uint32_t height);
/* sub-pixel offsets in [-0.5, 0.5) for PyrCamera.jitter: Halton (2, 3) over
 * 1024 frames, or a 4x4 grid visited in 16 frames, rows in bit-reversed
 * order and every other row backwards (reference pyrit.h:697-706) */
PYR_API void pyr_jitter_halton(uint32_t frame, float out[2]);
PYR_API void pyr_jitter_ordered(uint32_t frame, float out[2]);

/* ---- streamed worlds (reference pyrit.h:789-910) --------------------------
 *
 * A world keeps a LOD octree of chunks around the camera: every chunk has
 * 2^chunk_log2 cells per edge, and a cell of level `lod` spans 2^lod base
 * voxels. The application's generator writes each requested chunk's voxels on
 * the GPU; the chunks of one update are built as one GPU chunk batch and
 * entered into the context scene as instances (scale 2^lod, placed relative
 * to the render origin). A node is refined while its cells would look larger
 * than voxel_pixels on screen; children replace their parent only when all
 * eight are built (or empty). Generation and builds run on a worker thread
 * with its own stream unless PYR_WORLD_SYNC is set. Destroy a world before
 * its context. */

typedef struct PyrWorld PyrWorld;

typedef struct PyrChunkKey {
    int32_t x, y, z; /* chunk origin = (x, y, z) * 2^(chunk_log2 + lod) base voxels */
    uint32_t lod;
} PyrChunkKey;

/* One generator request. For every chunk c, write voxels (chunk-local cell
 * coordinates in [0, 2^chunk_log2), one cell = 2^lod base voxels) as
 *     uint32_t k = atomicAdd(&counts[c], 1);
 *     if (k < capacity) voxels[c * capacity + k] = v;
 * counts are zero on entry. Emit the visible skin, not full volumes. */
typedef struct PyrWorldGenParams {
    uint64_t chunks; /* device PyrChunkKey[count] */
    uint64_t voxels; /* device PyrVoxel[count * capacity] */
    uint64_t counts; /* device uint32_t[count] */
    uint32_t count, capacity, chunk_log2, reserved;
    uint64_t user;   /* PyrWorldInfo.user */
} PyrWorldGenParams;

/* Launches the application's kernels on `stream` (a hipStream_t); must not
 * synchronise. */
typedef void (*PyrWorldGenFn)(void* user, const PyrWorldGenParams* params, void* stream);

/* generate and build inside pyr_world_update / pyr_world_wait instead of on
 * the worker thread (deterministic; for tests and tools) */
#define PYR_WORLD_SYNC 0x1u

typedef struct PyrWorldInfo {
    uint32_t chunk_log2;        /* 3..8, 0 = 5 */
    uint32_t max_lod;           /* deepest level, 0 = 20 */
    uint32_t view_chunks;       /* radius of the coarsest level in chunks, 0 = 2 */
    float voxel_pixels;         /* target on-screen size of a cell, 0 = 4 */
    float view_distance;        /* base voxels, 0 = 16384 */
    uint32_t reserved0;
    uint64_t memory_budget;     /* bytes of chunk DAGs, 0 = 256 MiB */
    int32_t y_min, y_max;       /* vertical extent in base voxels (required) */
    uint32_t chunks_per_update; /* chunks per generated batch, 0 = chosen from the capacity */
    uint32_t mask;              /* instance mask, 0 = 0x1 */
    uint32_t secondary_mask;    /* instance mask bits of a second, coarser selection of the resident chunks
                                   (seen by shadow / GI rays); 0 = none */
    float secondary_pixels;     /* its target voxel size in pixels, 0 = 4 * voxel_pixels */
    uint32_t chunk_capacity;    /* generator slots per chunk, 0 = 8 * 4^chunk_log2 (at most 8^chunk_log2) */
    uint32_t rt_leaf_log2;      /* accepted (no RT cores on this device) */
    uint32_t keep_frames;       /* frames an unused chunk stays resident, 0 = 8 */
    uint32_t flags;             /* PYR_WORLD_* */
    PyrWorldGenFn generate;     /* required */
    void* user;
} PyrWorldInfo;

typedef struct PyrWorldStats {
    uint32_t visible_chunks, resident_chunks, pending_chunks;
    uint32_t built_chunks, built_voxels; /* in the last update */
    uint32_t overflow_chunks;            /* generator overflows (total): each grew the capacity per chunk and
                                            generated the job again; a chunk counts once more if it had to be
                                            clipped (the capacity could not grow) */
    float voxel_pixels;
    uint32_t top_lod;
    float secondary_bias;
    uint32_t path_edited;
    uint64_t bytes;                      /* device bytes of the resident chunk DAGs */
} PyrWorldStats;

/* one base voxel set (attribute != 0) or removed (attribute 0) */
typedef struct PyrWorldEdit {
    int64_t x, y, z;
    uint32_t attribute;
    uint32_t reserved;
} PyrWorldEdit;

PYR_API PyrResult pyr_world_create(PyrContext* ctx, const PyrWorldInfo* info, PyrWorld** out);
PYR_API PyrResult pyr_world_destroy(PyrContext* ctx, PyrWorld* world);
/* position: camera in base voxels; origin: render origin as in PyrFrameInfo
 * (NULL = 0); camera: next frame's camera for the screen-size LOD choice
 * (NULL = 1080p, 60 degrees). Never waits for the GPU (unless PYR_WORLD_SYNC). */
PYR_API PyrResult pyr_world_update(PyrContext* ctx, PyrWorld* world, const double position[3], const double origin[3],
                                   const PyrCamera* camera);
/* waits for the batch being generated and built, and takes it over */
PYR_API PyrResult pyr_world_wait(PyrContext* ctx, PyrWorld* world, const double camera[3]);
/* Edits overlay the generator: affected resident chunks are rebuilt at the
 * next update, and every later generation of a chunk applies them again. On
 * coarser levels a set fills its cell; a removal empties a cell only once all
 * of the cell's base voxels are removed. */
PYR_API PyrResult pyr_world_edit(PyrContext* ctx, PyrWorld* world, const PyrWorldEdit* edits, uint32_t count);
PYR_API uint64_t pyr_world_edits_bytes(PyrWorld* world);
PYR_API PyrResult pyr_world_edits_save(PyrWorld* world, void* dst, uint64_t size);
PYR_API PyrResult pyr_world_edits_load(PyrContext* ctx, PyrWorld* world, const void* src, uint64_t size);
PYR_API PyrResult pyr_world_stats(PyrWorld* world, PyrWorldStats* out);

#ifdef __cplusplus
}
#endif

#endif /* PYRIT_H */
