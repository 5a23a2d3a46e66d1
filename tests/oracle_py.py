"""ctypes access to the CPU oracle (oracle/liboracle.so) - the checker of the
parity tests. Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

u32, i32, u64, f32, f64 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double
P = C.POINTER

VOXEL = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("attribute", "<u4")])
RAY = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("t", "<f4"), ("instance", "<u4"), ("attribute", "<u4"), ("meta", "<u4")])
HIT_EX = np.dtype([("hit", HIT), ("voxel", "<i4", 3), ("reserved0", "<u4"), ("position", "<f4", 3),
                   ("reserved1", "<u4"), ("normal", "<f4", 3), ("reserved2", "<u4")])
RT_PRIM = np.dtype([("node", "<u4"), ("attr_base", "<u4"), ("cell", "<u8")])


class OrDag(C.Structure):
    _fields_ = [("log2_size", u32), ("root", u32), ("voxel_count", u64), ("nodes", P(u32)), ("node_words", u64),
                ("leaves", P(u64)), ("leaf_count", u64), ("attrs", P(u32)), ("attr_count", u64)]


class OrBuilt(C.Structure):
    _fields_ = [("log2_size", u32), ("root", u32), ("nodes", P(u32)), ("node_words", u32), ("leaves", P(u64)),
                ("leaf_count", u32), ("keys", P(u64)), ("attrs", P(u32)), ("voxel_count", u32)]


class OrDagHit(C.Structure):
    _fields_ = [("t", f32), ("face", u32), ("voxel", i32 * 3), ("attribute", u32), ("inside", u32)]


class OrDagView(C.Structure):
    _fields_ = [("nodes", P(u32)), ("leaves", P(u64)), ("attrs", P(u32)), ("root", u32), ("log2_size", u32),
                ("default_attribute", u32), ("attr_base", u32)]


class OrGeometry(C.Structure):
    _fields_ = [(n, u32) for n in ("node_offset", "leaf_offset", "attribute_offset", "root", "log2_size", "flags",
                                   "default_attribute", "palette_offset")]


class OrInstance(C.Structure):
    _fields_ = [("object_to_world", f32 * 12), ("world_to_object", f32 * 12), ("bounds_min", f32 * 3),
                ("geometry", u32), ("bounds_max", f32 * 3), ("mask", u32), ("user", u32), ("history", u32),
                ("flags", u32), ("reserved", u32)]


class OrScene(C.Structure):
    _fields_ = [("nodes", P(u32)), ("leaves", P(u64)), ("attrs", P(u32)), ("geometries", P(OrGeometry)),
                ("instances", P(OrInstance)), ("instance_count", u32), ("transparent", u64 * 4),
                ("cutout", u64 * 4)]


class OrCamera(C.Structure):
    _fields_ = [("view_to_world", f32 * 12), ("projection", u32), ("width", u32), ("height", u32),
                ("scale", f32 * 2), ("shift", f32 * 2), ("jitter", f32 * 2), ("near_plane", f32),
                ("far_plane", f32)]


class OrRangeAlloc(C.Structure):
    _fields_ = [("capacity", u64), ("used", u64), ("free", C.c_void_p), ("n", u32), ("cap", u32)]


_lib = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(
                os.path.join(ORACLE_DIR, "oracle.c")):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.or_morton.restype = u64
        L.or_morton.argtypes = [u32, u32, u32]
        L.or_morton_decode.argtypes = [u64, P(u32)]
        L.or_build_points.argtypes = [u32, C.c_void_p, C.c_size_t, i32, P(OrDag)]
        L.or_build_dense.argtypes = [u32, C.c_void_p, i32, P(OrDag)]
        L.or_dag_free.argtypes = [P(OrDag)]
        L.or_lookup.restype = u32
        L.or_lookup.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, u32, u32, u32, u32, u32]
        L.or_gb_build.argtypes = [u32, C.c_void_p, C.c_void_p, u32, C.c_void_p, u32, P(OrBuilt)]
        L.or_gb_build_chunks.argtypes = [u32, C.c_void_p, u32, u32, C.c_void_p, C.c_void_p, P(OrBuilt), C.c_void_p,
                                         C.c_void_p]
        L.or_built_free.argtypes = [P(OrBuilt)]
        L.or_canon.argtypes = [C.c_void_p, C.c_void_p, u32, u32, P(OrDag)]
        L.or_walk.argtypes = [P(OrDagView), P(f32), P(f32), f32, f32, i32, P(OrDagHit), P(u32)]
        L.or_walk_skipping.argtypes = [P(OrDagView), P(f32), P(f32), f32, f32, i32, C.c_void_p, C.c_void_p,
                                       P(i32), P(OrDagHit)]
        L.or_trace_rays.argtypes = [P(OrScene), C.c_void_p, C.c_void_p, u64, u32, u32, u32]
        L.or_trace_rays_ex.argtypes = [P(OrScene), C.c_void_p, C.c_void_p, u64, u32, u32, u32]
        L.or_trace_touch.argtypes = [P(OrScene), C.c_void_p, u64, u32, u32, P(u64)]
        L.or_camera_look_at.argtypes = [P(OrCamera), P(f32), P(f32), P(f32)]
        L.or_camera_perspective.argtypes = [P(OrCamera), f32, u32, u32, f32]
        L.or_camera_rays.argtypes = [P(OrCamera), C.c_void_p]
        L.or_inverse.argtypes = [P(f32), P(f32)]
        L.or_transformed_box.argtypes = [P(f32), f32, P(f32), P(f32)]
        L.or_render_primary.argtypes = [P(OrScene), P(OrCamera), C.c_void_p, C.c_void_p, u32]
        L.or_ref_dda.argtypes = [C.c_void_p, u32, P(f64), P(f64), f64, f64, P(f64), P(i32), P(i32)]
        L.or_range_init.argtypes = [P(OrRangeAlloc), u64]
        L.or_range_alloc.restype = C.c_int64
        L.or_range_alloc.argtypes = [P(OrRangeAlloc), u64]
        L.or_range_release.argtypes = [P(OrRangeAlloc), u64, u64]
        L.or_range_free.argtypes = [P(OrRangeAlloc)]
        L.or_palette_pack.argtypes = [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_rt_prims.restype = C.c_int64
        L.or_rt_prims.argtypes = [C.c_void_p, C.c_void_p, u32, u32, u32, C.c_void_p, C.c_void_p, u64, P(u32)]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None and len(a) else None


class Dag:
    """Host DAG (nodes, leaves, attrs, root, log2, voxel_count) as numpy."""

    def __init__(self, log2_size, root, nodes, leaves, attrs, voxel_count):
        self.log2_size, self.root = int(log2_size), int(root)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        self.leaves = np.ascontiguousarray(leaves, dtype=np.uint64)
        self.attrs = None if attrs is None else np.ascontiguousarray(attrs, dtype=np.uint32)
        self.voxel_count = int(voxel_count)

    @staticmethod
    def _from(d: OrDag) -> "Dag":
        nodes = np.ctypeslib.as_array(d.nodes, (d.node_words,)).copy() if d.node_words else np.zeros(0, np.uint32)
        leaves = np.ctypeslib.as_array(d.leaves, (d.leaf_count,)).copy() if d.leaf_count else np.zeros(0, np.uint64)
        attrs = None
        if d.attrs:
            attrs = np.ctypeslib.as_array(d.attrs, (d.attr_count,)).copy() if d.attr_count else np.zeros(0, np.uint32)
        return Dag(d.log2_size, d.root, nodes, leaves, attrs, d.voxel_count)

    def lookup(self, x, y, z) -> int:
        return lib().or_lookup(_p(self.nodes), _p(self.leaves), _p(self.attrs) if self.attrs is not None else None,
                               self.root, self.log2_size, x, y, z)

    def view(self) -> OrDagView:
        v = OrDagView()
        v.nodes = self.nodes.ctypes.data_as(P(u32))
        v.leaves = self.leaves.ctypes.data_as(P(u64)) if len(self.leaves) else None
        v.attrs = self.attrs.ctypes.data_as(P(u32)) if self.attrs is not None and len(self.attrs) else None
        v.root, v.log2_size, v.default_attribute, v.attr_base = self.root, self.log2_size, 1, 0
        return v

    def same(self, other: "Dag") -> bool:
        return (self.root == other.root and self.voxel_count == other.voxel_count
                and np.array_equal(self.nodes, other.nodes) and np.array_equal(self.leaves, other.leaves))


def build_points(log2_size: int, pts: np.ndarray, store_attrs: bool = True) -> Dag:
    d = OrDag()
    pts = np.ascontiguousarray(pts, dtype=VOXEL)
    rc = lib().or_build_points(log2_size, _p(pts), len(pts), int(store_attrs), C.byref(d))
    if rc != 0:
        raise ValueError(f"or_build_points failed ({rc})")
    try:
        return Dag._from(d)
    finally:
        lib().or_dag_free(C.byref(d))


def build_dense(log2_size: int, grid: np.ndarray, store_attrs: bool = True) -> Dag:
    d = OrDag()
    grid = np.ascontiguousarray(grid, dtype=np.uint32)
    rc = lib().or_build_dense(log2_size, _p(grid.ravel()), int(store_attrs), C.byref(d))
    if rc != 0:
        raise ValueError(f"or_build_dense failed ({rc})")
    try:
        return Dag._from(d)
    finally:
        lib().or_dag_free(C.byref(d))


def _built(b: OrBuilt) -> dict:
    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, (n,)).copy() if n else np.zeros(0, dt)

    return dict(log2_size=b.log2_size, root=b.root, nodes=arr(b.nodes, b.node_words, np.uint32),
                leaves=arr(b.leaves, b.leaf_count, np.uint64), keys=arr(b.keys, b.voxel_count, np.uint64),
                attrs=arr(b.attrs, b.voxel_count, np.uint32), voxel_count=b.voxel_count)


def gb_build(log2_size: int, edits: np.ndarray, old_keys=None, old_vals=None) -> dict:
    b = OrBuilt()
    edits = np.ascontiguousarray(edits, dtype=VOXEL)
    n_old = 0 if old_keys is None else len(old_keys)
    ok = None if old_keys is None else np.ascontiguousarray(old_keys, np.uint64)
    ov = None if old_vals is None else np.ascontiguousarray(old_vals, np.uint32)
    rc = lib().or_gb_build(log2_size, _p(ok), _p(ov), n_old, _p(edits), len(edits), C.byref(b))
    if rc != 0:
        raise ValueError(f"or_gb_build failed ({rc})")
    try:
        return _built(b)
    finally:
        lib().or_built_free(C.byref(b))


def gb_build_chunks(log2_size: int, voxels: np.ndarray, offsets: np.ndarray, chunk_log2=None, capacity: int = 0):
    """Packed (capacity 0) or segmented layout; returns (built, roots, first)."""
    b = OrBuilt()
    count = len(offsets) - 1
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    voxels = np.ascontiguousarray(voxels, dtype=VOXEL)
    cl = None if chunk_log2 is None else np.ascontiguousarray(chunk_log2, dtype=np.uint8)
    roots = np.zeros(count, np.uint32)
    first = np.zeros(count, np.uint32)
    rc = lib().or_gb_build_chunks(log2_size, _p(cl), count, capacity, _p(voxels), _p(offsets), C.byref(b),
                                  _p(roots), _p(first))
    if rc != 0:
        raise ValueError(f"or_gb_build_chunks failed ({rc})")
    try:
        return _built(b), roots, first
    finally:
        lib().or_built_free(C.byref(b))


def canon(nodes, leaves, root: int, log2_size: int) -> Dag:
    d = OrDag()
    nodes = np.ascontiguousarray(nodes, np.uint32)
    leaves = np.ascontiguousarray(leaves, np.uint64)
    rc = lib().or_canon(_p(nodes), _p(leaves) if len(leaves) else None, root, log2_size, C.byref(d))
    if rc != 0:
        raise ValueError("or_canon failed")
    try:
        return Dag._from(d)
    finally:
        lib().or_dag_free(C.byref(d))


def walk(dag: Dag, o, d, tmin=0.0, tmax=3.402823466e38, want_attr=True):
    h = OrDagHit()
    steps = u32(0)
    v = dag.view()
    of = (f32 * 3)(*[float(x) for x in o])
    df = (f32 * 3)(*[float(x) for x in d])
    hit = lib().or_walk(C.byref(v), of, df, tmin, tmax, int(want_attr), C.byref(h), C.byref(steps))
    return (h if hit else None), steps.value


def single_scene(dag: Dag, log2_size: int | None = None):
    """OrScene with one identity instance (bounds as xform.zig:37-58)."""
    L = lib()
    geo = (OrGeometry * 1)()
    geo[0].root, geo[0].log2_size = dag.root, dag.log2_size
    geo[0].flags = 1 if dag.attrs is not None else 0
    geo[0].default_attribute = 1
    inst = (OrInstance * 1)()
    ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    inst[0].object_to_world[:] = ident
    inst[0].world_to_object[:] = ident
    L.or_transformed_box(inst[0].object_to_world, float(1 << dag.log2_size), inst[0].bounds_min, inst[0].bounds_max)
    inst[0].mask, inst[0].flags = 0xFF, 1
    s = OrScene()
    s.nodes = dag.nodes.ctypes.data_as(P(u32))
    s.leaves = dag.leaves.ctypes.data_as(P(u64)) if len(dag.leaves) else None
    s.attrs = dag.attrs.ctypes.data_as(P(u32)) if dag.attrs is not None and len(dag.attrs) else None
    s.geometries = geo
    s.instances = inst
    s.instance_count = 1
    s._keep = (geo, inst, dag)
    return s


def multi_scene(dags: list, instances: list) -> OrScene:
    """OrScene over several DAGs (pools concatenated, geometry offsets into
    them) and instances given as (object_to_world[12], geometry index, mask);
    inverse and bounds as the product computes them (xform.zig:8-58)."""
    OL = lib()
    nodes = np.concatenate([d.nodes for d in dags]).astype(np.uint32)
    leaves = np.concatenate([d.leaves for d in dags] + [np.zeros(1, np.uint64)]).astype(np.uint64)
    attrs = np.concatenate([d.attrs for d in dags] + [np.zeros(1, np.uint32)]).astype(np.uint32)
    geo = (OrGeometry * len(dags))()
    no = lo = ao = 0
    for i, d in enumerate(dags):
        geo[i].node_offset, geo[i].leaf_offset, geo[i].attribute_offset = no, lo, ao
        geo[i].root, geo[i].log2_size, geo[i].flags, geo[i].default_attribute = d.root, d.log2_size, 1, 1
        no, lo, ao = no + len(d.nodes), lo + len(d.leaves), ao + len(d.attrs)
    inst = (OrInstance * len(instances))()
    for k, (m, g, mask) in enumerate(instances):
        inst[k].object_to_world[:] = [float(v) for v in m]
        OL.or_inverse(inst[k].object_to_world, inst[k].world_to_object)
        OL.or_transformed_box(inst[k].object_to_world, float(1 << dags[g].log2_size), inst[k].bounds_min,
                              inst[k].bounds_max)
        inst[k].geometry, inst[k].mask, inst[k].flags = g, mask, 1
    s = OrScene()
    s.nodes = nodes.ctypes.data_as(P(u32))
    s.leaves = leaves.ctypes.data_as(P(u64))
    s.attrs = attrs.ctypes.data_as(P(u32))
    s.geometries, s.instances, s.instance_count = geo, inst, len(instances)
    s._keep = (nodes, leaves, attrs, geo, inst, dags)
    return s


def trace_rays(scene: OrScene, rays: np.ndarray, ray_mask=0, flags=0, threads=8) -> np.ndarray:
    rays = np.ascontiguousarray(rays, dtype=RAY)
    hits = np.zeros(len(rays), HIT)
    lib().or_trace_rays(C.byref(scene), _p(rays), _p(hits), len(rays), ray_mask, flags, threads)
    return hits


def trace_rays_ex(scene: OrScene, rays: np.ndarray, ray_mask=0, flags=0, threads=8) -> np.ndarray:
    """PYR_TRACE_EXTENDED hits (trace.zig:125-157)"""
    rays = np.ascontiguousarray(rays, dtype=RAY)
    hits = np.zeros(len(rays), HIT_EX)
    lib().or_trace_rays_ex(C.byref(scene), _p(rays), _p(hits), len(rays), ray_mask, flags, threads)
    return hits


def trace_touch(scene: OrScene, rays: np.ndarray, ray_mask=0, flags=0) -> dict:
    """distinct node words, leaves and attributes read per ray, summed over the
    rays, and the hit count (one thread)"""
    rays = np.ascontiguousarray(rays, dtype=RAY)
    out = (C.c_uint64 * 4)()
    lib().or_trace_touch(C.byref(scene), _p(rays), len(rays), ray_mask, flags, out)
    return {"node_words": out[0], "leaves": out[1], "attributes": out[2], "hits": out[3], "rays": len(rays)}


def trace_dag_rays(dag: Dag, rays: np.ndarray, flags=0) -> np.ndarray:
    """walk per ray (instance 0), the semantics of pyrhip_trace_dag"""
    hits = np.zeros(len(rays), HIT)
    v = dag.view()
    h = OrDagHit()
    L = lib()
    for i, r in enumerate(rays):
        o = (f32 * 3)(*r["origin"].tolist())
        d = (f32 * 3)(*r["direction"].tolist())
        if L.or_walk(C.byref(v), o, d, float(r["tmin"]), float(r["tmax"]), int(not (flags & 2)), C.byref(h), None):
            hits[i] = (h.t, 0, h.attribute, h.face | (0x20 if h.inside else 0))
        else:
            hits[i] = (np.float32(3.402823466e38), 0xFFFFFFFF, 0, 0)
    return hits


def camera(eye, target, up, fov_y, width, height, near=0.0) -> OrCamera:
    c = OrCamera()
    L = lib()
    L.or_camera_look_at(C.byref(c), (f32 * 3)(*eye), (f32 * 3)(*target), (f32 * 3)(*up))
    L.or_camera_perspective(C.byref(c), fov_y, width, height, near)
    return c


def camera_rays(c: OrCamera) -> np.ndarray:
    rays = np.zeros(c.width * c.height, RAY)
    lib().or_camera_rays(C.byref(c), _p(rays))
    return rays


def shadow_rays(rays: np.ndarray, hits: np.ndarray, sun) -> tuple[np.ndarray, np.ndarray]:
    """Shadow ray per primary hit (shade.zig:626-629, 940): origin = hit point
    + face normal * 1e-3 (float32, evaluated in the reference's order),
    direction = sun, tmin 0, tmax FLT_MAX. Returns (rays, valid mask)."""
    valid = hits["instance"] != 0xFFFFFFFF
    face = hits["meta"] & 7
    axis = face >> 1
    sign = np.where(face & 1, np.float32(-1), np.float32(1)).astype(np.float32)
    out = np.zeros(len(rays), RAY)
    t = hits["t"].astype(np.float32)
    for a in range(3):
        nrm = np.where(axis == a, sign, np.float32(0)).astype(np.float32)
        p = (rays["origin"][:, a] + (rays["direction"][:, a] * t).astype(np.float32)).astype(np.float32)
        out["origin"][:, a] = (p + (nrm * np.float32(1e-3)).astype(np.float32)).astype(np.float32)
        out["direction"][:, a] = np.float32(sun[a])
    out["tmin"] = 0
    out["tmax"] = np.float32(3.402823466e38)
    return out, valid


def ref_dda(grid: np.ndarray, n: int, o, d, tmin=0.0, tmax=1e30):
    t = f64(0)
    v = (i32 * 3)()
    face = i32(0)
    hit = lib().or_ref_dda(_p(grid.ravel()), n, (f64 * 3)(*o), (f64 * 3)(*d), tmin, tmax, C.byref(t), v,
                           C.byref(face))
    return (t.value, tuple(v), face.value) if hit else None


def palette_pack(attrs: np.ndarray, first: np.ndarray):
    """(counts u32[k], word_offset u64[k+1], packed u32[total]) of or_palette_pack"""
    attrs = np.ascontiguousarray(attrs, np.uint32)
    first = np.ascontiguousarray(first, np.uint32)
    k = len(first) - 1
    counts = np.zeros(k, np.uint32)
    offs = np.zeros(k + 1, np.uint64)
    L = lib()
    L.or_palette_pack(_p(attrs), _p(first), k, _p(counts), _p(offs), None)
    out = np.zeros(max(int(offs[-1]), 1), np.uint32)
    L.or_palette_pack(_p(attrs), _p(first), k, _p(counts), _p(offs), _p(out))
    return counts, offs, out[:int(offs[-1])]


def rt_prims(dag: Dag, requested_log2: int):
    """(k, prims RT_PRIM[n], aabbs float32[n, 6]) of or_rt_prims"""
    L = lib()
    k = u32(0)
    leaves = _p(dag.leaves)
    n = L.or_rt_prims(_p(dag.nodes), leaves, dag.root, dag.log2_size, requested_log2, None, None, 0, C.byref(k))
    if n < 0:
        raise MemoryError("or_rt_prims")
    prims = np.zeros(max(n, 1), RT_PRIM)
    aabbs = np.zeros((max(n, 1), 6), np.float32)
    L.or_rt_prims(_p(dag.nodes), leaves, dag.root, dag.log2_size, requested_log2, _p(prims), _p(aabbs.ravel()), n,
                  C.byref(k))
    return k.value, prims[:n], aabbs[:n]
