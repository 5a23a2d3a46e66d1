"""Python mirror of the reference's build / trace operator interface
(src/gpu_build.zig:94-131 build / buildChunks, src/device/trace.zig trace)
over the C ABI of libpyrit.so. Device memory comes from pyrit_amd.device."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi, lib
from .device import Buffer, download, to_device


@dataclass
class Built:
    """gpu_build.Built (src/gpu_build.zig:25-47): device arrays owned here."""
    raw: abi.PyrhipBuilt

    @property
    def log2_size(self): return self.raw.log2_size
    @property
    def root(self): return self.raw.root
    @property
    def node_words(self): return self.raw.node_words
    @property
    def leaf_count(self): return self.raw.leaf_count
    @property
    def voxel_count(self): return self.raw.voxel_count

    def nodes(self) -> np.ndarray:
        return download(self.raw.nodes, np.uint32, self.raw.node_words)

    def leaves(self) -> np.ndarray:
        return download(self.raw.leaves, np.uint64, self.raw.leaf_count)

    def attrs(self) -> np.ndarray:
        return download(self.raw.attrs, np.uint32, self.raw.voxel_count)

    def keys(self) -> np.ndarray:
        return download(self.raw.keys, np.uint64, self.raw.voxel_count)

    def dag(self) -> abi.PyrhipDag:
        d = abi.PyrhipDag()
        d.nodes, d.leaves, d.attrs = self.raw.nodes, self.raw.leaves, self.raw.attrs
        d.root, d.log2_size, d.default_attribute = self.raw.root, self.raw.log2_size, 1
        return d

    def free(self) -> None:
        if self.raw.nodes or self.raw.keys:
            lib().pyrhip_built_free(C.byref(self.raw), None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _tuning(radix_bits: int, sort_items: int = 0, pair_layout: int = 0, scan_items: int = 0,
            hash_load_pct: int = 0, status_words: int = 0, finish: int = 0,
            counted_pass: int = 0, levels: int = 0) -> abi.PyrhipBuildTuning:
    """0 = the setting installed for the workload class, else the default;
    pair_layout 1 = separate key / attribute arrays, 2 = packed when keys fit"""
    t = abi.PyrhipBuildTuning()
    t.radix_bits, t.sort_items, t.pair_layout, t.scan_items = radix_bits, sort_items, pair_layout, scan_items
    t.hash_load_pct = hash_load_pct
    t.status_words = status_words
    t.finish = finish
    t.counted_pass = counted_pass
    t.levels = levels
    return t


def _ptr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, Buffer):
        return x.ptr
    return int(x)


def build(log2_size: int, voxels, count: int | None = None, old: Built | None = None, radix_bits: int = 0,
          stream: int | None = None, sort_items: int = 0, pair_layout: int = 0, scan_items: int = 0,
          hash_load_pct: int = 0, status_words: int = 0, finish: int = 0, counted_pass: int = 0,
          levels: int = 0) -> Built:
    """Build a DAG from PyrVoxel records (numpy VOXEL array or device Buffer).
    `old` supplies the sorted voxels of an earlier editable build (edits)."""
    keep = None
    if isinstance(voxels, np.ndarray):
        assert voxels.dtype == abi.VOXEL
        count = len(voxels)
        keep = to_device(voxels) if count else None
        vptr = keep.ptr if keep else 0
    else:
        vptr = _ptr(voxels)
        assert count is not None
    out = abi.PyrhipBuilt()
    t = _tuning(radix_bits, sort_items, pair_layout, scan_items, hash_load_pct, status_words, finish, counted_pass,
                levels)
    L = lib()
    code = L.pyrhip_build(log2_size, old.raw.keys if old else 0, old.raw.attrs if old else 0,
                          old.voxel_count if old else 0, vptr, count, C.byref(t), stream, C.byref(out))
    abi.check(L, code, "pyrhip_build")
    return Built(out)


def _batch(voxels, offsets: np.ndarray, chunk_log2: np.ndarray | None, capacity: int):
    """PyrhipChunkBatch of host or device voxels; the second value keeps the
    device copies alive"""
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    count = len(offsets) - 1
    keep = []
    if isinstance(voxels, np.ndarray):
        vb = to_device(voxels) if len(voxels) else Buffer(16)
        keep.append(vb)
        vptr = vb.ptr
    else:
        vptr = _ptr(voxels)
    ob = to_device(offsets)
    keep.append(ob)
    lptr = 0
    if chunk_log2 is not None:
        lb = to_device(np.ascontiguousarray(chunk_log2, dtype=np.uint8))
        keep.append(lb)
        lptr = lb.ptr
    batch = abi.PyrhipChunkBatch()
    batch.count, batch.capacity, batch.voxels, batch.offsets = count, capacity, vptr, ob.ptr
    batch.total, batch.chunk_log2 = int(offsets[-1]), lptr
    return batch, keep


def build_chunks(log2_size: int, voxels, offsets: np.ndarray, chunk_log2: np.ndarray | None = None,
                 capacity: int = 0, radix_bits: int = 0, stream: int | None = None, hash_load_pct: int = 0,
                 status_words: int = 0, finish: int = 0, counted_pass: int = 0, levels: int = 0):
    """Chunk batch (src/gpu_build.zig:98-122). voxels: VOXEL array or device
    Buffer; offsets: u32[count+1] host array. Returns (Built, roots, first)."""
    batch, keep = _batch(voxels, offsets, chunk_log2, capacity)
    count = batch.count
    roots = np.zeros(count, np.uint32)
    first = np.zeros(count, np.uint32)
    out = abi.PyrhipBuilt()
    t = _tuning(radix_bits, hash_load_pct=hash_load_pct, status_words=status_words, finish=finish,
                counted_pass=counted_pass, levels=levels)
    L = lib()
    code = L.pyrhip_build_chunks(log2_size, C.byref(batch), C.byref(t), stream, C.byref(out),
                                 roots.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 first.ctypes.data_as(C.POINTER(C.c_uint32)))
    abi.check(L, code, "pyrhip_build_chunks")
    return Built(out), roots, first


def trace_dag(dag: abi.PyrhipDag, rays, flags: int = 0, stream: int | None = None) -> np.ndarray:
    """PyrRay records -> PyrHit records against one DAG (instance 0)."""
    keep = None
    if isinstance(rays, np.ndarray):
        assert rays.dtype == abi.RAY
        count = len(rays)
        keep = to_device(rays)
        rptr = keep.ptr
    else:
        rptr, count = rays
    hits = Buffer(max(count, 1) * abi.HIT.itemsize)
    L = lib()
    abi.check(L, L.pyrhip_trace_dag(C.byref(dag), rptr, hits.ptr, count, flags, stream), "pyrhip_trace_dag")
    return hits.download(abi.HIT, count)


def trace_camera(dag: abi.PyrhipDag, cam: abi.PyrCamera, flags: int = 0, stream: int | None = None) -> np.ndarray:
    n = cam.width * cam.height
    hits = Buffer(max(n, 1) * abi.HIT.itemsize)
    L = lib()
    abi.check(L, L.pyrhip_trace_camera(C.byref(dag), C.byref(cam), hits.ptr, flags, stream), "pyrhip_trace_camera")
    return hits.download(abi.HIT, n)


def camera_rays(cam: abi.PyrCamera, stream: int | None = None) -> np.ndarray:
    n = cam.width * cam.height
    rays = Buffer(max(n, 1) * abi.RAY.itemsize)
    L = lib()
    abi.check(L, L.pyrhip_camera_rays(C.byref(cam), rays.ptr, stream), "pyrhip_camera_rays")
    return rays.download(abi.RAY, n)


def trace_shadow(dag: abi.PyrhipDag, rays: np.ndarray, hits: np.ndarray, sun, stream: int | None = None) -> np.ndarray:
    count = len(rays)
    rb, hb = to_device(rays), to_device(hits)
    out = Buffer(max(count, 1) * 4)
    s = (C.c_float * 3)(*[float(v) for v in sun])
    L = lib()
    abi.check(L, L.pyrhip_trace_shadow(C.byref(dag), rb.ptr, hb.ptr, s, out.ptr, count, stream),
              "pyrhip_trace_shadow")
    return out.download(np.uint32, count)


def rt_prims(dag: abi.PyrhipDag, node_words: int, leaf_count: int, requested_log2: int,
             stream: int | None = None):
    """Box primitives of a DAG (src/rt_prims.zig:127-137 extract):
    (k, prims RT_PRIM[n], aabbs float32[n, 6])."""
    L = lib()
    count, k = C.c_uint32(0), C.c_uint32(0)
    abi.check(L, L.pyrhip_rt_prims(C.byref(dag), node_words, leaf_count, requested_log2, 0, 0, 0, stream,
                                   C.byref(count), C.byref(k)), "pyrhip_rt_prims")
    n = count.value
    prims = Buffer(max(n, 1) * abi.RT_PRIM.itemsize)
    aabbs = Buffer(max(n, 1) * 24)
    abi.check(L, L.pyrhip_rt_prims(C.byref(dag), node_words, leaf_count, requested_log2, prims.ptr, aabbs.ptr, n,
                                   stream, C.byref(count), C.byref(k)), "pyrhip_rt_prims")
    return k.value, prims.download(abi.RT_PRIM, n), aabbs.download(np.float32, 6 * n).reshape(n, 6)


def set_timing(on: bool) -> None:
    lib().pyrhip_set_timing(1 if on else 0)


def last_build_times() -> dict:
    t = abi.PyrhipBuildTimes()
    lib().pyrhip_last_build_times(C.byref(t))
    return {k: getattr(t, k) for k, _ in abi.PyrhipBuildTimes._fields_}


def _tuning_dict(t: abi.PyrhipBuildTuning) -> dict:
    return {k: int(getattr(t, k)) for k, _ in abi.PyrhipBuildTuning._fields_}


def autotune(log2_size: int, voxels, count: int | None = None, offsets: np.ndarray | None = None,
             chunk_log2: np.ndarray | None = None, repeats: int = 3, stream: int | None = None) -> dict:
    """pyrhip_autotune on a single build (offsets None) or a chunk batch;
    installs the winner for the workload class and returns the report with
    the class key ({"log2_size", "n", "chunk_count"}) for tuning_set."""
    keep = []
    batch = None
    if offsets is not None:
        batch, keep = _batch(voxels, offsets, chunk_log2, 0)
        vptr, n = 0, batch.total
    elif isinstance(voxels, np.ndarray):
        vb = to_device(voxels) if len(voxels) else Buffer(16)
        keep.append(vb)
        vptr, n = vb.ptr, len(voxels)
    else:
        vptr, n = _ptr(voxels), int(count)
    rep = abi.PyrhipTuneReport()
    L = lib()
    abi.check(L, L.pyrhip_autotune(log2_size, vptr, n, C.byref(batch) if batch is not None else None, repeats,
                                   stream, C.byref(rep)), "pyrhip_autotune")
    return {"class": {"log2_size": log2_size, "n": n, "chunk_count": batch.count if batch is not None else 0},
            "best": _tuning_dict(rep.best), "best_ms": rep.best_ms, "default_ms": rep.default_ms,
            "candidates": rep.candidates, "rejected": rep.rejected}


def tuning_get(log2_size: int, n: int, chunk_count: int = 0) -> dict:
    t = abi.PyrhipBuildTuning()
    L = lib()
    abi.check(L, L.pyrhip_tuning_get(log2_size, n, chunk_count, C.byref(t)), "pyrhip_tuning_get")
    return _tuning_dict(t)


def tuning_set(log2_size: int, n: int, chunk_count: int, best: dict | None) -> None:
    """install a setting (e.g. one saved from an autotune report) or remove it (None)"""
    L = lib()
    t = _tuning(**best) if best is not None else None
    abi.check(L, L.pyrhip_tuning_set(log2_size, n, chunk_count, C.byref(t) if t is not None else None),
              "pyrhip_tuning_set")


def tuning_clear() -> None:
    lib().pyrhip_tuning_clear()


class GraphOutput(Built):
    """A graph launch's output: arrays in the graph's buffers, valid until its
    next launch; never freed here."""

    def free(self) -> None:
        pass


class BuildGraph:
    """A chunk-batch build captured once into a HIP graph (pyrhip_graph_*),
    replayed for any batch of at most `count` chunks and `total_max` voxels."""

    def __init__(self, log2_size: int, count: int, total_max: int, capacity: int = 0, mixed_sizes: bool = False,
                 finish: int = 0):
        self.count = count
        self.handle = C.c_void_p()
        L = lib()
        t = _tuning(0, finish=finish)
        abi.check(L, L.pyrhip_graph_create(log2_size, count, capacity, total_max, int(mixed_sizes), C.byref(t),
                                           C.byref(self.handle)), "pyrhip_graph_create")

    def voxels_ptr(self) -> int:
        return lib().pyrhip_graph_voxels(self.handle)

    def launch(self, voxels, offsets: np.ndarray, chunk_log2: np.ndarray | None = None, stream: int | None = None):
        """-> (GraphOutput, roots, first); offsets: u32[k + 1] for k <= count
        chunks (the rest are empty); voxels: VOXEL array, device Buffer /
        address, or None when they are already in voxels_ptr()"""
        offs = np.zeros(self.count + 1, np.uint32)
        o = np.ascontiguousarray(offsets, np.uint32)
        offs[:len(o)] = o
        offs[len(o):] = o[-1]  # trailing chunks empty
        keep = None
        if isinstance(voxels, np.ndarray):
            assert voxels.dtype == abi.VOXEL
            keep = to_device(voxels) if len(voxels) else None
            vptr = keep.ptr if keep else 0
        else:
            vptr = _ptr(voxels)
        l2 = None
        if chunk_log2 is not None:
            l2 = np.full(self.count, 3, np.uint8)
            l2[:len(chunk_log2)] = chunk_log2
        out = abi.PyrhipBuilt()
        roots = np.zeros(self.count, np.uint32)
        first = np.zeros(self.count, np.uint32)
        L = lib()
        abi.check(L, L.pyrhip_graph_launch(self.handle, vptr, offs.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           l2.ctypes.data_as(C.POINTER(C.c_uint8)) if l2 is not None else None,
                                           stream, C.byref(out), roots.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           first.ctypes.data_as(C.POINTER(C.c_uint32))), "pyrhip_graph_launch")
        return GraphOutput(out), roots, first

    def destroy(self) -> None:
        if self.handle:
            lib().pyrhip_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
