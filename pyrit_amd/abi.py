"""ctypes mirrors of the C ABI (include/pyrit.h, include/pyrit_hip_build.h,
pyrit_amd/csrc/terrain.h) and numpy dtypes of the record arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

u8, u32, i32, u64, f32, f64 = C.c_uint8, C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double
vp = C.c_void_p

# ---- record dtypes (byte-identical to the C structs) ----------------------
VOXEL = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("attribute", "<u4")])
RAY = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("t", "<f4"), ("instance", "<u4"), ("attribute", "<u4"), ("meta", "<u4")])
HIT_EX = np.dtype([("hit", HIT), ("voxel", "<i4", 3), ("reserved0", "<u4"), ("position", "<f4", 3),
                   ("reserved1", "<u4"), ("normal", "<f4", 3), ("reserved2", "<u4")])  # PyrHitEx
RT_PRIM = np.dtype([("node", "<u4"), ("attr_base", "<u4"), ("cell", "<u8")])  # PyrhipRtPrim

PYR_OK = 0
PYR_ERROR_INVALID_ARGUMENT = -1
PYR_ERROR_INVALID_HANDLE = -2
PYR_ERROR_OUT_OF_MEMORY = -3
PYR_ERROR_CAPACITY = -4
PYR_ERROR_IN_USE = -5
PYR_ERROR_CUDA = -6
PYR_ERROR_COMPILE = -7
PYR_ERROR_NOT_FOUND = -8
PYR_ERROR_VERSION = -9
PYR_FLT_MAX = np.float32(3.402823466e38)
PYR_NO_HIT = 0xFFFFFFFF
PYR_HIT_INSIDE = 0x20
PYR_PROJECTION_PERSPECTIVE = 0
PYR_PROJECTION_ORTHOGRAPHIC = 1
PYR_TRACE_ANY_HIT = 0x1
PYR_TRACE_NO_ATTRIBUTE = 0x2
PYR_TRACE_EXTENDED = 0x4
PYR_TRACE_SKIP_TRANSPARENT = 0x8
PYR_BUILD_HOST_INPUT = 0x1
PYR_BUILD_EDITABLE = 0x2
PYR_GEOMETRY_HAS_ATTRIBUTES = 0x1
PYR_GEOMETRY_PALETTE = 0x2
PYR_DAG_NO_ATTRIBUTES = 0x1
PYR_VERSION = (0 << 16) | 2


class PyrCamera(C.Structure):
    _fields_ = [("view_to_world", f32 * 12), ("projection", u32), ("width", u32), ("height", u32),
                ("scale", f32 * 2), ("shift", f32 * 2), ("jitter", f32 * 2), ("near_plane", f32),
                ("far_plane", f32)]


class PyrGeometryData(C.Structure):
    _fields_ = [("node_offset", u32), ("leaf_offset", u32), ("attribute_offset", u32), ("root", u32),
                ("log2_size", u32), ("flags", u32), ("default_attribute", u32), ("palette_offset", u32)]


class PyrInstanceData(C.Structure):
    _fields_ = [("object_to_world", f32 * 12), ("world_to_object", f32 * 12), ("bounds_min", f32 * 3),
                ("geometry", u32), ("bounds_max", f32 * 3), ("mask", u32), ("user", u32), ("history", u32),
                ("flags", u32), ("reserved", u32)]


class PyrScene(C.Structure):
    _fields_ = [("nodes", u64), ("leaves", u64), ("attributes", u64), ("geometries", u64), ("instances", u64),
                ("instances_prev", u64), ("instance_count", u32), ("geometry_count", u32), ("frame", u64),
                ("time", f64), ("time_prev", f64), ("origin", f64 * 3), ("origin_prev", f64 * 3),
                ("materials", u64), ("lighting", u64), ("transparent_materials", u64 * 4),
                ("cutout_materials", u64 * 4), ("textures", u64), ("texture_count", u32), ("reserved_tex", u32)]


class PyrCreateInfo(C.Structure):
    _fields_ = [("struct_size", u32), ("version", u32), ("device", i32), ("flags", u32), ("cuda_context", vp),
                ("cuda_stream", vp), ("max_instances", u32), ("max_geometries", u32), ("max_views", u32),
                ("max_textures", u32), ("rt_leaf_log2", u32), ("node_pool_bytes", u64), ("leaf_pool_bytes", u64),
                ("attribute_pool_bytes", u64), ("staging_bytes", u64), ("log", vp), ("log_user", vp)]


class PyrFrameInfo(C.Structure):
    _fields_ = [("time", f64), ("origin", f64 * 3)]


class PyrMaterial(C.Structure):
    _fields_ = [("base_color", f32 * 3), ("roughness", f32), ("emission", f32 * 3), ("metallic", f32),
                ("flags", u32), ("opacity", f32), ("ior", f32), ("density", f32), ("wave_height", f32),
                ("wave_length", f32), ("wave_speed", f32), ("clearcoat", f32), ("clearcoat_roughness", f32),
                ("subsurface", f32), ("subsurface_color", f32 * 3), ("texture", u32), ("normal_texture", u32),
                ("texture_scale", f32), ("normal_strength", f32), ("normal_scale", f32), ("side_texture", u32),
                ("side_color", f32 * 3), ("variation", f32 * 3), ("variation_scale", f32 * 2), ("variation_warm", f32)]


PYR_MAX_MATERIALS = 256
PYR_MATERIAL_VOXEL_COLOR = 0x1
PYR_MATERIAL_TRANSPARENT = 0x4
PYR_MATERIAL_CUTOUT = 0x10


class PyrTargets(C.Structure):
    _fields_ = [("hits", u64), ("depth", u64), ("motion", u64), ("color", u64), ("normal", u64), ("albedo", u64),
                ("material", u64), ("ray_mask", u32), ("flags", u32), ("transparent_mask", u32),
                ("secondary_mask", u32), ("coverage", u32), ("detail_scale", f32)]


class PyrhipBuilt(C.Structure):
    _fields_ = [("log2_size", u32), ("root", u32), ("nodes", u64), ("node_words", u32), ("leaf_count", u32),
                ("leaves", u64), ("keys", u64), ("attrs", u64), ("voxel_count", u32), ("reserved", u32)]


class PyrhipChunkBatch(C.Structure):
    _fields_ = [("count", u32), ("capacity", u32), ("voxels", u64), ("offsets", u64), ("total", u32),
                ("reserved", u32), ("chunk_log2", u64)]


class PyrhipBuildTuning(C.Structure):
    _fields_ = [("radix_bits", u32), ("sort_items", u32), ("pair_layout", u32), ("scan_items", u32),
                ("hash_load_pct", u32), ("status_words", u32), ("finish", u32), ("counted_pass", u32),
                ("levels", u32)]


class PyrhipTuneReport(C.Structure):
    _fields_ = [("best", PyrhipBuildTuning), ("best_ms", f32), ("default_ms", f32), ("candidates", u32),
                ("rejected", u32)]


class PyrhipBuildTimes(C.Structure):
    _fields_ = [("encode_ms", f32), ("sort_ms", f32), ("unique_ms", f32), ("leaves_ms", f32), ("levels_ms", f32),
                ("total_ms", f32), ("sort_passes", u32), ("radix_bits", u32), ("record_bytes", u32),
                ("sort_items", u32), ("host_alloc_ms", f32), ("host_enqueue_ms", f32), ("host_wait_ms", f32),
                ("status_bytes", u32), ("finish", u32), ("device_bytes", u64)]


class PyrhipBuildOutput(C.Structure):
    _fields_ = [("nodes", u64), ("node_capacity", u64), ("leaves", u64), ("leaf_capacity", u64), ("keys", u64),
                ("attrs", u64), ("voxel_capacity", u64)]


class PyrhipDag(C.Structure):
    _fields_ = [("nodes", u64), ("leaves", u64), ("attrs", u64), ("root", u32), ("log2_size", u32),
                ("default_attribute", u32), ("reserved", u32), ("palette", u64)]


class PyrhipServerInfo(C.Structure):
    _fields_ = [("memory_budget", u64), ("max_batch_chunks", u32), ("max_batch_voxels", u32), ("workers", u32),
                ("flags", u32), ("tuning", PyrhipBuildTuning)]


class PyrhipRequest(C.Structure):
    _fields_ = [("chunk", u64), ("kind", u32), ("log2_size", u32), ("voxels", u64), ("count", u32),
                ("reserved", u32), ("user", u64)]


class PyrhipCompletion(C.Structure):
    _fields_ = [("ticket", u64), ("chunk", u64), ("user", u64), ("result", i32), ("voxel_count", u32),
                ("submit_s", f64), ("start_s", f64), ("done_s", f64), ("batch_chunks", u32), ("reserved", u32)]


class PyrhipServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("rejected", u64), ("batches", u64),
                ("resident_chunks", u64), ("resident_bytes", u64), ("reserved_bytes", u64), ("pending", u64),
                ("batched_voxels", u64), ("busy_s", f64)]


PYRHIP_REQUEST_BUILD = 0
PYRHIP_REQUEST_EDIT = 1
PYRHIP_SERVER_NO_GRAPHS = 1


class PyrChunkKey(C.Structure):
    _fields_ = [("x", i32), ("y", i32), ("z", i32), ("lod", u32)]


class PyrWorldGenParams(C.Structure):
    _fields_ = [("chunks", u64), ("voxels", u64), ("counts", u64), ("count", u32), ("capacity", u32),
                ("chunk_log2", u32), ("reserved", u32), ("user", u64)]


class PyrWorldInfo(C.Structure):
    _fields_ = [("chunk_log2", u32), ("max_lod", u32), ("view_chunks", u32), ("voxel_pixels", f32),
                ("view_distance", f32), ("reserved0", u32), ("memory_budget", u64), ("y_min", i32), ("y_max", i32),
                ("chunks_per_update", u32), ("mask", u32), ("secondary_mask", u32), ("secondary_pixels", f32),
                ("chunk_capacity", u32), ("rt_leaf_log2", u32), ("keep_frames", u32), ("flags", u32),
                ("generate", vp), ("user", vp)]


class PyrWorldStats(C.Structure):
    _fields_ = [("visible_chunks", u32), ("resident_chunks", u32), ("pending_chunks", u32), ("built_chunks", u32),
                ("built_voxels", u32), ("overflow_chunks", u32), ("voxel_pixels", f32), ("top_lod", u32),
                ("secondary_bias", f32), ("path_edited", u32), ("bytes", u64)]


class PyrWorldEdit(C.Structure):
    _fields_ = [("x", C.c_int64), ("y", C.c_int64), ("z", C.c_int64), ("attribute", u32), ("reserved", u32)]


PYR_WORLD_SYNC = 0x1


class PyrTerrainWorld(C.Structure):
    _fields_ = [("seed", u32), ("mean", i32), ("amplitude", i32), ("octaves", u32), ("wavelength_log2", u32),
                ("reserved", u32)]


class PyrTerrainCave(C.Structure):
    _fields_ = [("x", i32), ("y", i32), ("z", i32), ("r", i32)]


class PyrTerrainChunk(C.Structure):
    _fields_ = [("origin", i32 * 3), ("log2_size", u32), ("seed", u32), ("mean", i32), ("amplitude", i32),
                ("octaves", u32), ("wavelength_log2", u32), ("material_mode", u32), ("skin", u32),
                ("water_level", i32), ("cave_first", u32), ("cave_count", u32), ("grid_log2", u32),
                ("grid_dim", u32), ("grid_first", u32)]


class PyrStats(C.Structure):
    _fields_ = [("node_pool_used", u64), ("node_pool_capacity", u64), ("leaf_pool_used", u64),
                ("leaf_pool_capacity", u64), ("attribute_pool_used", u64), ("attribute_pool_capacity", u64),
                ("geometries", u32), ("instances", u32), ("frame", u64), ("features", u32), ("reserved", u32)]


class PyrDagInfo(C.Structure):
    _fields_ = [("log2_size", u32), ("root", u32), ("voxel_count", u64), ("nodes", C.POINTER(u32)),
                ("node_words", u64), ("leaves", C.POINTER(u64)), ("leaf_count", u64),
                ("attributes", C.POINTER(u32)), ("attribute_count", u64)]


P = C.POINTER
VOXEL_FN = C.CFUNCTYPE(u32, vp, i32, i32, i32)
REGION_EMPTY_FN = C.CFUNCTYPE(i32, vp, i32, i32, i32, u32)

# name -> (restype, argtypes) of every C entry point of libpyrit.so
FUNCTIONS = {
    "pyr_result_string": (C.c_char_p, [i32]),
    "pyr_error_message": (C.c_char_p, []),
    "pyr_create": (i32, [P(PyrCreateInfo), P(vp)]),
    "pyr_destroy": (None, [vp]),
    "pyr_cuda_stream": (vp, [vp]),
    "pyr_features": (u32, [vp]),
    "pyr_synchronize": (i32, [vp]),
    "pyr_geometry_build": (i32, [vp, u32, vp, u32, u32, P(vp)]),
    "pyr_geometry_edit": (i32, [vp, vp, vp, u32, u32]),
    "pyr_geometry_create": (i32, [vp, vp, P(vp)]),
    "pyr_geometry_downsample": (i32, [vp, vp, u32, u32, P(vp)]),
    "pyr_geometry_download": (i32, [vp, vp, P(vp)]),
    "pyr_geometry_destroy": (i32, [vp, vp]),
    "pyr_geometry_index": (u32, [vp]),
    "pyr_instance_create": (i32, [vp, vp, P(vp)]),
    "pyr_instance_destroy": (i32, [vp, vp]),
    "pyr_instance_set_transform": (i32, [vp, vp, P(f32)]),
    "pyr_instance_set_mask": (i32, [vp, vp, u32]),
    "pyr_instance_set_geometry": (i32, [vp, vp, vp]),
    "pyr_get_stats": (i32, [vp, P(PyrStats)]),
    "pyr_instance_set_user": (i32, [vp, vp, u32]),
    "pyr_instance_index": (u32, [vp]),
    "pyr_commit": (i32, [vp, P(PyrFrameInfo)]),
    "pyr_scene_device": (u64, [vp]),
    "pyr_view_create": (i32, [vp, P(vp)]),
    "pyr_view_destroy": (i32, [vp, vp]),
    "pyr_view_reset_history": (i32, [vp, vp]),
    "pyr_instance_reset_history": (i32, [vp, vp]),
    "pyr_jitter_halton": (None, [u32, P(f32)]),
    "pyr_jitter_ordered": (None, [u32, P(f32)]),
    "pyr_render": (i32, [vp, vp, P(PyrCamera), P(PyrTargets)]),
    "pyr_trace": (i32, [vp, u64, u64, u32, u32, u32]),
    "pyr_material_default": (None, [P(PyrMaterial)]),
    "pyr_material_set": (i32, [vp, u32, P(PyrMaterial)]),
    "pyr_voxel_attribute": (u32, [u32, u32, u32, u32]),
    "pyr_camera_look_at": (None, [P(PyrCamera), P(f32), P(f32), P(f32)]),
    "pyr_camera_perspective": (None, [P(PyrCamera), f32, u32, u32, f32]),
    "pyr_camera_orthographic": (None, [P(PyrCamera), f32, u32, u32]),
    "pyrhip_build": (i32, [u32, u64, u64, u32, u64, u32, P(PyrhipBuildTuning), vp, P(PyrhipBuilt)]),
    "pyrhip_build_static": (i32, [u32, u64, u64, u32, u64, u32, P(PyrhipBuildTuning), vp, P(PyrhipBuildOutput),
                                  P(PyrhipBuilt)]),
    "pyrhip_build_static_limit": (u64, []),
    "pyrhip_build_chunks": (i32, [u32, P(PyrhipChunkBatch), P(PyrhipBuildTuning), vp, P(PyrhipBuilt), P(u32),
                                  P(u32)]),
    "pyrhip_built_free": (None, [P(PyrhipBuilt), vp]),
    "pyrhip_graph_create": (i32, [u32, u32, u32, u32, i32, P(PyrhipBuildTuning), P(vp)]),
    "pyrhip_graph_voxels": (u64, [vp]),
    "pyrhip_graph_launch": (i32, [vp, u64, P(u32), P(u8), vp, P(PyrhipBuilt), P(u32), P(u32)]),
    "pyrhip_graph_destroy": (None, [vp]),
    "pyrhip_decode_voxels": (i32, [u64, u64, u32, u32, u64, vp]),
    "pyrhip_autotune": (i32, [u32, u64, u32, P(PyrhipChunkBatch), u32, vp, P(PyrhipTuneReport)]),
    "pyrhip_tuning_set": (i32, [u32, u32, u32, P(PyrhipBuildTuning)]),
    "pyrhip_tuning_get": (i32, [u32, u32, u32, P(PyrhipBuildTuning)]),
    "pyrhip_tuning_clear": (None, []),
    "pyrhip_path_edit": (i32, [P(PyrhipDag), u32, u32, u32, u32, u32, P(u32), u32, u64, vp, P(u32), P(u32), P(u32)]),
    "pyrhip_palette_pack": (i32, [u64, P(u32), u32, u64, u64, P(u64), P(u32), vp]),
    "pyrhip_rt_prims": (i32, [P(PyrhipDag), u32, u32, u32, u64, u64, u32, vp, P(u32), P(u32)]),
    "pyrhip_downsample": (i32, [u64, u64, u32, u32, u32, P(PyrhipBuildTuning), vp, P(PyrhipBuilt)]),
    "pyrhip_set_timing": (None, [i32]),
    "pyrhip_copy_probe": (i32, [u64, u64, u64, vp]),
    "pyrhip_last_build_times": (None, [P(PyrhipBuildTimes)]),
    "pyrhip_pool_stats": (i32, [P(u64)]),
    "pyrhip_trim": (None, []),
    "pyrhip_trace_dag": (i32, [P(PyrhipDag), u64, u64, u32, u32, vp]),
    "pyrhip_trace_camera": (i32, [P(PyrhipDag), P(PyrCamera), u64, u32, vp]),
    "pyrhip_camera_rays": (i32, [P(PyrCamera), u64, vp]),
    "pyrhip_trace_shadow": (i32, [P(PyrhipDag), u64, u64, P(f32), u64, u32, vp]),
    "pyrhip_server_create": (i32, [P(PyrhipServerInfo), P(vp)]),
    "pyrhip_server_submit": (i32, [vp, P(PyrhipRequest), P(u64)]),
    "pyrhip_server_poll": (u32, [vp, P(PyrhipCompletion), u32]),
    "pyrhip_server_drain": (i32, [vp]),
    "pyrhip_server_chunk": (i32, [vp, u64, P(PyrhipDag), P(u32), P(u32), P(u32)]),
    "pyrhip_server_clock": (f64, []),
    "pyrhip_server_stats": (None, [vp, P(PyrhipServerStats)]),
    "pyrhip_server_error": (u32, [vp, C.c_char_p, u32]),
    "pyrhip_server_destroy": (None, [vp]),
    "pyr_world_create": (i32, [vp, P(PyrWorldInfo), P(vp)]),
    "pyr_world_destroy": (i32, [vp, vp]),
    "pyr_world_update": (i32, [vp, vp, P(f64), P(f64), P(PyrCamera)]),
    "pyr_world_wait": (i32, [vp, vp, P(f64)]),
    "pyr_world_edit": (i32, [vp, vp, vp, u32]),
    "pyr_world_edits_bytes": (u64, [vp]),
    "pyr_world_edits_save": (i32, [vp, vp, u64]),
    "pyr_world_edits_load": (i32, [vp, vp, vp, u64]),
    "pyr_world_stats": (i32, [vp, P(PyrWorldStats)]),
    "pyr_dag_build_dense": (i32, [u32, vp, u32, P(vp)]),
    "pyr_dag_build_points": (i32, [u32, vp, C.c_size_t, u32, P(vp)]),
    "pyr_dag_build_fn": (i32, [u32, VOXEL_FN, REGION_EMPTY_FN, vp, u32, P(vp)]),
    "pyr_dag_get_info": (None, [vp, P(PyrDagInfo)]),
    "pyr_dag_destroy": (None, [vp]),
    "pyr_dag_save": (i32, [vp, vp, P(C.c_size_t)]),
    "pyr_dag_load": (i32, [vp, C.c_size_t, P(vp)]),
}

TERRAIN_FUNCTIONS = {
    "pyrgen_generate_host": (i32, [P(PyrTerrainChunk), u32, P(PyrTerrainCave), P(u32), P(u32), P(u64), vp]),
    "pyrgen_scratch_bytes": (u64, [u64]),
    "pyrgen_height_host": (i32, [P(PyrTerrainChunk), i32, i32]),
    "pyrgen_generate_device": (i32, [u64, u32, u64, u64, u64, u64, u64, u64, u64, u64, vp]),
    "pyrgen_world_generate": (None, [vp, P(PyrWorldGenParams), vp]),
    "pyrgen_world_chunk_host": (u32, [P(PyrTerrainWorld), P(PyrChunkKey), u32, vp, u32]),
}


def declare(lib: C.CDLL) -> None:
    for name, (res, args) in FUNCTIONS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def declare_terrain(lib: C.CDLL) -> None:
    for name, (res, args) in TERRAIN_FUNCTIONS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def check(lib: C.CDLL, code: int, what: str = "") -> None:
    if code != PYR_OK:
        msg = lib.pyr_error_message().decode(errors="replace")
        name = lib.pyr_result_string(code).decode()
        raise RuntimeError(f"{what}: {name} ({code}): {msg}")
