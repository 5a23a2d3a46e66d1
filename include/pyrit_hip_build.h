/*
 * pyrit_hip_build.h - the internal operator boundary of the GPU DAG build and
 * the software traversal, as C-ABI entry points over device pointers.
 *
 * These replace the reference's build-side `gpu_build.Exec` vtable and its
 * entry points (src/gpu_build.zig:14-23, 94-131) and the traversal launches
 * pyr_k_trace / pyr_k_render (src/gpu_kernels.zig:63-85). Every kernel behind
 * them is hand-written HIP for gfx950 (libpyrit.so).
 *
 * Conventions: device pointers are passed as uint64_t, streams as void*
 * (hipStream_t). All work is enqueued on `stream`; calls that return sizes
 * synchronise that stream once at the end (the reference reads its counters
 * back once, src/gpu_build.zig:338-344). Results are PYR_* codes (pyrit.h).
 */
#ifndef PYRIT_HIP_BUILD_H
#define PYRIT_HIP_BUILD_H

#include <stdint.h>

#include "pyrit.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Built DAG in the GPU-native layout (src/gpu_build.zig:25-47): nodes level by
 * level bottom-up, word refs absolute, leaves in first-appearance Morton order,
 * attributes in DFS rank order. All buffers are device memory owned by the
 * caller, taken from the library's device-memory cache: release them with
 * pyrhip_built_free, not hipFree. `keys`/`attrs` hold the sorted unique
 * voxels (brick Morton << 6 | bit) for later edits. */
typedef struct PyrhipBuilt {
    uint32_t log2_size;
    uint32_t root;
    uint64_t nodes;
    uint32_t node_words;
    uint32_t leaf_count;
    uint64_t leaves;
    uint64_t keys;
    uint64_t attrs;
    uint32_t voxel_count;
    uint32_t reserved;
} PyrhipBuilt;

/* One chunk batch (src/gpu_build.zig:98-117). Chunk c's voxels (PyrVoxel,
 * chunk-local coordinates) are voxels[c*capacity + j] for
 * j < offsets[c+1]-offsets[c] when capacity > 0 (the generator layout of
 * pyrit.h:819-834), or voxels[offsets[c] .. offsets[c+1]) when capacity == 0
 * (packed). offsets is a device array of count+1 u32, total = offsets[count].
 * chunk_log2 (device u8[count], may be 0 = all log2_size) gives each chunk its
 * own edge 2^chunk_log2[c] <= 2^log2_size: chunks of mixed LOD sizes build in
 * one batch and each gets its root at its own top level. */
typedef struct PyrhipChunkBatch {
    uint32_t count;
    uint32_t capacity;
    uint64_t voxels;
    uint64_t offsets;
    uint32_t total;
    uint32_t reserved;
    uint64_t chunk_log2;
} PyrhipChunkBatch;

/* Tuning of one build. Every setting gives the same output; they trade
 * passes, tile sizes (the work of one 256-thread block), LDS footprint,
 * occupancy and hash probe lengths. A field left 0 takes the value installed
 * for the build's workload class by pyrhip_autotune / pyrhip_tuning_set, else
 * the built-in default. */
typedef struct PyrhipBuildTuning {
    uint32_t radix_bits;   /* digit width of the LSD passes: 8 (default), 9, 10 or 11 */
    uint32_t sort_items;   /* keys per thread of a sort tile (256 threads): 8, 12, 16, 20 (default), 24 */
    uint32_t pair_layout;  /* 2 (default): pack key << 32 | attribute when keys fit 32 bits (8-bit digits);
                              1: always separate key and attribute arrays */
    uint32_t scan_items;   /* records per thread of a unique / brick scan tile: 8, 12 (default), 16, 20 */
    uint32_t hash_load_pct; /* maximum load of the leaf and node hash tables in percent, 25..95 (default 80);
                               the output does not depend on it */
    uint32_t status_words;  /* sort look-back words: 2 (default) 32 bits when a build of packed records holds
                               fewer than 2^30 of them, 1 always 64 bits */
    uint32_t finish;        /* 2: sort only the key bits above the low 14 and take the unique voxels and
                               bricks from a table per bucket of 2^14 keys of a chunk; 1: full sort and unique
                               scan; 0 (default): 2 when a single build has 2048 records per bucket on average
                               (a batch 512 per bucket of chunks of its largest size) */
    uint32_t counted_pass;  /* 1 (default): with 8-bit digits the first radix pass takes every tile's output
                               offsets from per-tile digit counts the encode step writes (later passes use the
                               decoupled look-back); 2: decoupled look-back in every pass */
    uint32_t levels;        /* node levels: 2 one workgroup builds every level in one launch, 1 four launches
                               per level over the whole GPU, 0 (default): 2 when the build holds at most
                               32768 bricks (its brick bound), else 1; the output is the same */
} PyrhipBuildTuning;

/* gpu_build.build (src/gpu_build.zig:94-96): DAG of n_old sorted unique
 * voxels (old_keys/old_vals of an earlier build, may be 0) plus n_new PyrVoxel
 * records at `voxels`; later records win, attribute 0 deletes, out-of-range
 * records are dropped. log2_size in [3, 20]. */
PYR_API PyrResult pyrhip_build(uint32_t log2_size, uint64_t old_keys, uint64_t old_vals, uint32_t n_old,
                               uint64_t voxels, uint32_t n_new, const PyrhipBuildTuning* tuning, void* stream,
                               PyrhipBuilt* out);

/* Caller-owned device buffers that pyrhip_build_static writes its results
 * into: node words, leaves, and (optional, 0 = not wanted) the sorted unique
 * voxel keys and their attributes, with their capacities in elements. */
typedef struct PyrhipBuildOutput {
    uint64_t nodes;
    uint64_t node_capacity;
    uint64_t leaves;
    uint64_t leaf_capacity;
    uint64_t keys;
    uint64_t attrs;
    uint64_t voxel_capacity;
} PyrhipBuildOutput;

/* pyrhip_build without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Every
 * scratch buffer of the build comes from a static device array of
 * pyrhip_build_static_limit() bytes (64 MiB) that each device holds until the
 * process ends; the results are copied into `dst` (out points at dst's
 * buffers). A build whose scratch does not fit fails with PYR_ERROR_CAPACITY
 * and a message naming the bytes it needs; so does output larger than dst's
 * capacities. Calls on one device share that array: calls that overlap on
 * one GPU (several host threads, or streams that run concurrently) race. */
PYR_API PyrResult pyrhip_build_static(uint32_t log2_size, uint64_t old_keys, uint64_t old_vals, uint32_t n_old,
                                      uint64_t voxels, uint32_t n_new, const PyrhipBuildTuning* tuning, void* stream,
                                      const PyrhipBuildOutput* dst, PyrhipBuilt* out);
PYR_API uint64_t pyrhip_build_static_limit(void);

/* gpu_build.buildChunks (src/gpu_build.zig:119-122). All chunks share the
 * node, leaf and attribute arrays; roots[c] / first_voxel[c] (host arrays of
 * count) get each chunk's root word and first attribute rank, 0xFFFFFFFF for
 * an empty chunk. */
PYR_API PyrResult pyrhip_build_chunks(uint32_t log2_size, const PyrhipChunkBatch* batch,
                                      const PyrhipBuildTuning* tuning, void* stream, PyrhipBuilt* out,
                                      uint32_t* roots, uint32_t* first_voxel);

PYR_API void pyrhip_built_free(PyrhipBuilt* built, void* stream);

/* ---- chunk-batch builds captured into a HIP graph (no counterpart in the
 * reference, which launches every kernel of gpu_build.zig:181-336 one by
 * one) ---------------------------------------------------------------------
 * The whole build of a batch of `count` chunks holding up to `total_max`
 * voxels (capacity layout when capacity > 0, packed otherwise; per-chunk
 * edges given at launch when mixed_sizes != 0) is captured once, with every
 * buffer allocated once; a launch uploads the offsets, copies the voxels
 * into the graph's input buffer (skipped when they are already there, see
 * pyrhip_graph_voxels), replays the graph and reads the counters back. The
 * kernels take their sizes from the device offsets, so any batch of at most
 * `count` chunks (trailing chunks empty) and `total_max` voxels replays the
 * same graph. The outputs live in the graph's buffers: valid until the next
 * launch or pyrhip_graph_destroy, never passed to pyrhip_built_free. One
 * launch at a time per graph. */
typedef struct PyrhipGraph PyrhipGraph;
PYR_API PyrResult pyrhip_graph_create(uint32_t log2_size, uint32_t count, uint32_t capacity, uint32_t total_max,
                                      int32_t mixed_sizes, const PyrhipBuildTuning* tuning, PyrhipGraph** out);
/* device address of the graph's voxel input buffer */
PYR_API uint64_t pyrhip_graph_voxels(PyrhipGraph* graph);
/* offsets: host u32[count + 1]; chunk_log2: host u8[count] for mixed sizes, else NULL; voxels: device
 * (0 = already in the input buffer). The replay runs on `stream` and has finished when the call returns. */
PYR_API PyrResult pyrhip_graph_launch(PyrhipGraph* graph, uint64_t voxels, const uint32_t* offsets,
                                      const uint8_t* chunk_log2, void* stream, PyrhipBuilt* out, uint32_t* roots,
                                      uint32_t* first_voxel);
PYR_API void pyrhip_graph_destroy(PyrhipGraph* graph);

/* gpu_build.downsample (src/gpu_build.zig:124-131): DAG of edge
 * 2^(log2_src - shift) from n_src sorted unique voxels (src_keys/src_vals of
 * an editable build); every source voxel moves to its coarse cell and, per
 * cell, the voxel with the largest source key wins. shift >= 1 and
 * log2_src - shift >= 3. */
PYR_API PyrResult pyrhip_downsample(uint64_t src_keys, uint64_t src_vals, uint32_t n_src, uint32_t log2_src,
                                    uint32_t shift, const PyrhipBuildTuning* tuning, void* stream, PyrhipBuilt* out);

/* ---- tuning per workload class (no counterpart in the reference, whose
 * build has fixed launch shapes: gpu_build.zig:186-190) -------------------
 * A class is (log2_size, bit length of the record count, bit length of the
 * chunk count; 0 for single builds). */

typedef struct PyrhipTuneReport {
    PyrhipBuildTuning best; /* every field explicit */
    float best_ms;          /* median build time of the winner (HIP events on the stream) */
    float default_ms;       /* the same with the built-in defaults */
    uint32_t candidates;    /* settings measured */
    uint32_t rejected;      /* settings that failed or whose output differed from the defaults' (0 expected) */
} PyrhipTuneReport;

/* Measures settings on this workload (a single build of n PyrVoxel records at
 * `voxels`, or `batch` when not NULL), each `repeats` times (0: 3) after one
 * untimed run, by coordinate descent over sort_items, scan_items, radix_bits /
 * pair_layout, hash_load_pct and status_words; a change must be 2 % faster to be taken. The
 * winner is installed for the workload's class. Synchronises `stream`. */
PYR_API PyrResult pyrhip_autotune(uint32_t log2_size, uint64_t voxels, uint32_t n, const PyrhipChunkBatch* batch,
                                  uint32_t repeats, void* stream, PyrhipTuneReport* out);
/* installs (tuning != NULL) or removes the setting of a class, e.g. one saved
 * from an earlier autotune run */
PYR_API PyrResult pyrhip_tuning_set(uint32_t log2_size, uint32_t n, uint32_t chunk_count,
                                    const PyrhipBuildTuning* tuning);
/* the setting a build of this class uses when its tuning fields are 0 */
PYR_API PyrResult pyrhip_tuning_get(uint32_t log2_size, uint32_t n, uint32_t chunk_count, PyrhipBuildTuning* out);
PYR_API void pyrhip_tuning_clear(void);

/* Palette packing of a batch's attributes (the reference's
 * pyr_k_palette_scan / pyr_k_palette_pack, src/device/palette.zig, used by
 * context.zig:1311-1370 when a world installs a batch). Chunk j owns ranks
 * [first[j], first[j+1]) of `attrs` (first: host u32[count + 1]). On return
 * palette_count[j] (host u32[count]) is its number of distinct attributes, or
 * 0xFFFFFFFF when there are more than 16 (or the value 0xFFFFFFFF occurs), and
 * word_offset[j] (host u64[count + 1]; the last entry is the total) its start
 * in the output: 16 palette entries in ascending order (padded with the
 * first) and then 4-bit indices, eight per word (rank r in bits 4 (r & 7) of
 * word r >> 3); or its 32-bit attributes. out == 0 computes the layout only;
 * otherwise out (device, out_words words) receives the packed attributes. */
PYR_API PyrResult pyrhip_palette_pack(uint64_t attrs, const uint32_t* first, uint32_t count, uint64_t out,
                                      uint64_t out_words, uint64_t* word_offset, uint32_t* palette_count,
                                      void* stream);

/* Sorted voxel keys of a build (Built.keys/attrs, or a chunk's slice of a
 * batch) back to PyrVoxel records in chunk-local coordinates; log2_size is the
 * edge the keys were built with (the batch's for a batch slice). The input of
 * an edit rebuild of a chunk of a batch (the reference keeps old keys for the
 * same purpose, src/gpu_build.zig:90-96). */
PYR_API PyrResult pyrhip_decode_voxels(uint64_t keys, uint64_t attrs, uint32_t count, uint32_t log2_size,
                                       uint64_t out, void* stream);

/* Last build's kernel timings (device ms, from HIP events on the build
 * stream) when timing was enabled with pyrhip_set_timing(1). */
typedef struct PyrhipBuildTimes {
    float encode_ms, sort_ms, unique_ms, leaves_ms, levels_ms, total_ms;
    uint32_t sort_passes, radix_bits;
    uint32_t record_bytes; /* bytes per record moved by a sort pass: 8 packed, 12 pairs */
    uint32_t sort_items;
    /* host milliseconds: allocating device memory, enqueueing the kernels,
     * waiting for the final read-back */
    float host_alloc_ms, host_enqueue_ms, host_wait_ms;
    uint32_t status_bytes; /* bytes per sort look-back word: 4 or 8 */
    uint32_t finish;       /* 2: partial sort + bucket tables, 1: full sort + unique scan */
    uint64_t device_bytes; /* device memory the build requested, scratch and outputs (all held at once:
                              the peak) */
} PyrhipBuildTimes;
PYR_API void pyrhip_set_timing(int enable);
/* Plain device-to-device copy kernel of `bytes` (16-byte aligned buffers,
 * a multiple of 16 bytes): the calibration probe for HBM byte counters
 * (reads = writes = bytes) and the copy ceiling of the device. No
 * counterpart in the reference. */
PYR_API PyrResult pyrhip_copy_probe(uint64_t src, uint64_t dst, uint64_t bytes, void* stream);
PYR_API void pyrhip_last_build_times(PyrhipBuildTimes* out);
/* The device-memory cache that builds, their outputs and server generations
 * allocate from (freed blocks stay mapped for reuse): out = {limit of cached
 * free bytes, bytes held, bytes in use, bytes cached free}. pyrhip_trim gives
 * every cached free block back to the driver. No counterpart in the
 * reference (its builds allocate from the CUDA stream-ordered pool). */
PYR_API PyrResult pyrhip_pool_stats(uint64_t out[4]);
PYR_API void pyrhip_trim(void);

/* One DAG in device memory for pyrhip_trace_dag (a single instance with the
 * identity transform; src/device/dag.zig:24-36 Dag). */
typedef struct PyrhipDag {
    uint64_t nodes;
    uint64_t leaves;
    uint64_t attrs;        /* 0: every hit reports default_attribute */
    uint32_t root;
    uint32_t log2_size;
    uint32_t default_attribute;
    uint32_t reserved;
    uint64_t palette;      /* 0, or 16 attribute values: attrs then holds 4-bit indices, eight per word */
} PyrhipDag;

/* Path edit (the reference's pyr_k_path_edit, src/device/pathedit.zig): sets
 * or removes voxels (x, y, z, attribute; 0 removes) of a finished DAG without
 * a rebuild. Only the paths from the root to the edited bricks are written
 * again: a new brick goes to leaves[leaf_count ..] and the new nodes to
 * nodes[node_words ..], up to the given capacities (device arrays; the old
 * words stay valid). attrs_out (device, voxel_count + count words) receives
 * the new attribute array in DFS rank order; dag->attrs is read only. Edits
 * apply in the order given. out_flags: 0 = applied (out_root, out_count are
 * the new root and voxel count), 2 = the DAG became empty, 4 = the capacities
 * were too small (nothing usable was written). */
PYR_API PyrResult pyrhip_path_edit(const PyrhipDag* dag, uint32_t voxel_count, uint32_t node_words,
                                   uint32_t node_capacity, uint32_t leaf_count, uint32_t leaf_capacity,
                                   const uint32_t* edits, uint32_t count, uint64_t attrs_out, void* stream,
                                   uint32_t* out_root, uint32_t* out_count, uint32_t* out_flags);

/* One box primitive of a DAG (src/device/types.zig:518-525 RtPrim): the
 * subtree root `node` (word offset), the DFS attribute rank of its first
 * voxel, and its cell in units of the primitive edge, x | y << 21 | z << 42
 * (src/device/rt.zig:16-18 packCell). 16 bytes. */
typedef struct PyrhipRtPrim {
    uint32_t node;
    uint32_t attr_base;
    uint64_t cell;
} PyrhipRtPrim;

/* Box primitives of a finished DAG (src/rt_prims.zig:127-137 extract; the
 * same list the reference's GPU build emits at rt_log2, src/device/
 * gbuild.zig:574-605, 251-257): every non-empty subtree of edge 2^k,
 * k = clamp(requested_log2, 3, log2_size) (*out_log2), in depth-first order,
 * with its voxel box (max exclusive) widened by 1/256: aabbs[6 i ..] = min
 * x, y, z, max x, y, z as float (the OptixAabb layout, rt_prims.zig:14).
 * node_words / leaf_count bound every reference the call reads. prims and
 * aabbs (device, capacity entries) may both be 0: then only *out_count is
 * set. More primitives than capacity: PYR_ERROR_CAPACITY with *out_count the
 * number needed, nothing written. For a chunk batch, call once per chunk root
 * and append: the running count before a chunk is its first primitive
 * (gbuild.zig:251-257 prim_start). */
PYR_API PyrResult pyrhip_rt_prims(const PyrhipDag* dag, uint32_t node_words, uint32_t leaf_count,
                                  uint32_t requested_log2, uint64_t prims, uint64_t aabbs, uint32_t capacity,
                                  void* stream, uint32_t* out_count, uint32_t* out_log2);

/* PyrRay[count] -> PyrHit[count] (device pointers) against one DAG, the
 * traversal of src/device/dag.zig:160-351 plus toHit (src/device/trace.zig:
 * 115-118) with instance 0. flags: PYR_TRACE_ANY_HIT / PYR_TRACE_NO_ATTRIBUTE. */
PYR_API PyrResult pyrhip_trace_dag(const PyrhipDag* dag, uint64_t rays, uint64_t hits, uint32_t count,
                                   uint32_t flags, void* stream);

/* Primary rays of a camera (src/device/render.zig:16-32, 133-141) traced
 * against one DAG: hits[width*height], rays generated on the device. */
PYR_API PyrResult pyrhip_trace_camera(const PyrhipDag* dag, const PyrCamera* camera, uint64_t hits,
                                      uint32_t flags, void* stream);

/* The primary rays of a camera, PyrRay[width*height] row-major (device). */
PYR_API PyrResult pyrhip_camera_rays(const PyrCamera* camera, uint64_t rays, void* stream);

/* Shadow rays: for every hit of a primary pass (rays[i], hits[i]), a ray from
 * the hit point offset 1e-3 along the face normal toward `sun` (normalised),
 * any-hit (src/device/shade.zig:626-629, 940). out[i] = 1 occluded, 0 lit,
 * 0xFFFFFFFF for a primary miss. */
PYR_API PyrResult pyrhip_trace_shadow(const PyrhipDag* dag, uint64_t rays, uint64_t hits, const float sun[3],
                                      uint64_t out, uint32_t count, void* stream);

/* ---------------------------------------------------------------------------
 * Serving: a continuous-batching build server with resident chunk DAGs.
 *
 * The reference streams world chunks through a job pipeline on a worker
 * thread that generates, applies edits and builds one batch at a time
 * (src/world.zig:755-1124: submitJob / workerMain / gpuWorkImpl / finishJob,
 * with capacity growth). Here requests arrive one by one: worker threads (one
 * HIP stream each) repeatedly take every pending request that fits the device
 * memory budget, build them as one mixed-size chunk batch, keep each chunk's
 * DAG and sorted voxels resident, and complete the requests. Requests that
 * arrive while batches are on the device join the next batch; with two or
 * more workers batches overlap on the device. Requests on the same chunk are
 * applied in submission order. Results are identical to building each
 * request alone (the batch output is canonicalised per chunk by construction:
 * same voxels, same canonical DAG).
 * ------------------------------------------------------------------------- */

typedef struct PyrhipServer PyrhipServer;

typedef struct PyrhipServerInfo {
    uint64_t memory_budget;     /* device bytes for resident DAGs plus batches in flight; 0 = 64 GiB */
    uint32_t max_batch_chunks;  /* 0 = 1024 */
    uint32_t max_batch_voxels;  /* records per batch, 0 = 2^28 */
    uint32_t workers;           /* batches in flight (one stream each), 0 = 2 */
    uint32_t flags;             /* PYRHIP_SERVER_* */
    PyrhipBuildTuning tuning;
} PyrhipServerInfo;

/* Build every batch with pyrhip_build_chunks. Without it, a batch of at most
 * 64 chunks, 2^20 records and edges up to 2^8 replays a HIP graph
 * (pyrhip_graph_*, one per worker and largest edge 2^3..2^8, up to about 100 MB each, captured by
 * pyrhip_server_create and kept until pyrhip_server_destroy; not counted in
 * memory_budget). No capture runs while the server serves: HIP fails
 * legacy-stream calls of every thread of the process during a capture. */
#define PYRHIP_SERVER_NO_GRAPHS 1u

#define PYRHIP_REQUEST_BUILD 0u /* (re)build a chunk from voxel records (device memory) */
#define PYRHIP_REQUEST_EDIT 1u  /* set / remove voxels of a resident chunk (host records, copied) */

typedef struct PyrhipRequest {
    uint64_t chunk;      /* chunk id */
    uint32_t kind;       /* PYRHIP_REQUEST_* */
    uint32_t log2_size;  /* chunk edge (build; edit of a chunk that is not resident) */
    uint64_t voxels;     /* build: device PyrVoxel*; edit: host PyrVoxel* */
    uint32_t count;
    uint32_t reserved;
    uint64_t user;
} PyrhipRequest;

typedef struct PyrhipCompletion {
    uint64_t ticket, chunk, user;
    int32_t result;          /* PYR_OK or the error of the batch / admission */
    uint32_t voxel_count;    /* voxels of the chunk after the request */
    double submit_s, start_s, done_s; /* steady clock: submitted, batch launched, DAG resident */
    uint32_t batch_chunks;   /* size of the batch that carried the request */
    uint32_t reserved;
} PyrhipCompletion;

typedef struct PyrhipServerStats {
    uint64_t submitted, completed, rejected, batches;
    uint64_t resident_chunks, resident_bytes, reserved_bytes, pending;
    uint64_t batched_voxels;
    double busy_s; /* summed device time of the batches (HIP events) */
} PyrhipServerStats;

PYR_API PyrResult pyrhip_server_create(const PyrhipServerInfo* info, PyrhipServer** out);
PYR_API PyrResult pyrhip_server_submit(PyrhipServer* server, const PyrhipRequest* request, uint64_t* ticket);
/* completed requests since the last poll (at most max); non-blocking */
PYR_API uint32_t pyrhip_server_poll(PyrhipServer* server, PyrhipCompletion* out, uint32_t max);
/* block until every request submitted so far has completed */
PYR_API PyrResult pyrhip_server_drain(PyrhipServer* server);
/* resident DAG of a chunk, valid until the chunk is rebuilt or the server is
 * destroyed: the node and leaf arrays of the batch that built it (shared with
 * the batch's other chunks; node_words / leaf_count give their lengths), the
 * chunk's root and its attributes. An empty chunk has nodes == 0. */
PYR_API PyrResult pyrhip_server_chunk(PyrhipServer* server, uint64_t chunk, PyrhipDag* out, uint32_t* voxel_count,
                                      uint32_t* node_words, uint32_t* leaf_count);
PYR_API double pyrhip_server_clock(void); /* the steady clock of the completion time stamps */
PYR_API void pyrhip_server_stats(PyrhipServer* server, PyrhipServerStats* out);
/* message of the last failed batch or rejected request ("" if none) into out
 * (size bytes, truncated); returns the message's length. A worker's batch
 * fails on its own thread, so pyr_error_message of the caller does not hold it. */
PYR_API uint32_t pyrhip_server_error(PyrhipServer* server, char* out, uint32_t size);
PYR_API void pyrhip_server_destroy(PyrhipServer* server);

#ifdef __cplusplus
}
#endif
#endif
