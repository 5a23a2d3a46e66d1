"""Host DAGs through the C ABI (pyr_dag_*, pyr_geometry_create/download):
numpy views of a PyrDag and the binary PYRD format (the reference's
src/io.zig:86-160)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, lib


class HostDag:
    """Owns one PyrDag*; arrays() copies its node, leaf and attribute arrays."""

    def __init__(self, handle: C.c_void_p):
        self.handle = handle

    # ---- constructors ------------------------------------------------------
    @staticmethod
    def build_points(log2_size: int, voxels: np.ndarray, flags: int = 0) -> "HostDag":
        v = np.ascontiguousarray(voxels, dtype=abi.VOXEL)
        out = C.c_void_p()
        L = lib()
        abi.check(L, L.pyr_dag_build_points(log2_size, v.ctypes.data_as(C.c_void_p) if len(v) else None, len(v),
                                            flags, C.byref(out)), "pyr_dag_build_points")
        return HostDag(out)

    @staticmethod
    def build_dense(log2_size: int, grid: np.ndarray, flags: int = 0) -> "HostDag":
        """grid indexed [z, y, x], i.e. voxels[x + N * (y + N * z)] in C order"""
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        out = C.c_void_p()
        L = lib()
        abi.check(L, L.pyr_dag_build_dense(log2_size, g.ctypes.data_as(C.c_void_p), flags, C.byref(out)),
                  "pyr_dag_build_dense")
        return HostDag(out)

    @staticmethod
    def build_fn(log2_size: int, voxel, region_empty=None, flags: int = 0) -> "HostDag":
        """voxel(x, y, z) -> attribute, region_empty(x, y, z, size) -> bool"""
        fn = abi.VOXEL_FN(lambda _u, x, y, z: int(voxel(x, y, z)))
        re = abi.REGION_EMPTY_FN(lambda _u, x, y, z, s: int(bool(region_empty(x, y, z, s)))) if region_empty \
            else abi.REGION_EMPTY_FN()
        out = C.c_void_p()
        L = lib()
        abi.check(L, L.pyr_dag_build_fn(log2_size, fn, re, None, flags, C.byref(out)), "pyr_dag_build_fn")
        return HostDag(out)

    @staticmethod
    def load(data: bytes) -> "HostDag":
        out = C.c_void_p()
        L = lib()
        buf = C.create_string_buffer(bytes(data), len(data))
        abi.check(L, L.pyr_dag_load(buf, len(data), C.byref(out)), "pyr_dag_load")
        return HostDag(out)

    @staticmethod
    def download(ctx: C.c_void_p, geometry: C.c_void_p) -> "HostDag":
        out = C.c_void_p()
        L = lib()
        abi.check(L, L.pyr_geometry_download(ctx, geometry, C.byref(out)), "pyr_geometry_download")
        return HostDag(out)

    # ---- access ------------------------------------------------------------
    def info(self) -> abi.PyrDagInfo:
        i = abi.PyrDagInfo()
        lib().pyr_dag_get_info(self.handle, C.byref(i))
        return i

    def arrays(self) -> dict:
        i = self.info()
        nodes = np.ctypeslib.as_array(i.nodes, (i.node_words,)).copy() if i.node_words else np.zeros(0, np.uint32)
        leaves = np.ctypeslib.as_array(i.leaves, (i.leaf_count,)).copy() if i.leaf_count else np.zeros(0, np.uint64)
        attrs = None
        if i.attributes:
            attrs = np.ctypeslib.as_array(i.attributes, (i.attribute_count,)).copy() if i.attribute_count \
                else np.zeros(0, np.uint32)
        return {"log2_size": i.log2_size, "root": i.root, "voxel_count": i.voxel_count, "nodes": nodes,
                "leaves": leaves, "attrs": attrs}

    def save(self) -> bytes:
        L = lib()
        n = C.c_size_t(0)
        abi.check(L, L.pyr_dag_save(self.handle, None, C.byref(n)), "pyr_dag_save size")
        buf = C.create_string_buffer(n.value)
        abi.check(L, L.pyr_dag_save(self.handle, buf, C.byref(n)), "pyr_dag_save")
        return buf.raw[:n.value]

    def create_geometry(self, ctx: C.c_void_p) -> C.c_void_p:
        g = C.c_void_p()
        L = lib()
        abi.check(L, L.pyr_geometry_create(ctx, self.handle, C.byref(g)), "pyr_geometry_create")
        return g

    def destroy(self) -> None:
        if self.handle:
            lib().pyr_dag_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
