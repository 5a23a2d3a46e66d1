"""ABI checks in the style of the reference's tests/abi_test.zig:11-55: the
library loads (no GPU needed), exports every function include/*.h declares,
and the struct layouts of the C header match the Python mirrors and the
reference's sizes."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("pyrit.h", "pyrit_hip_build.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
        for m in re.finditer(r"PYR_API\s+[^;(]*?\b(\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_declared_symbols():
    from pyrit_amd import LIB_PATH, abi

    assert os.path.exists(LIB_PATH), "libpyrit.so not built"
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    # every declared function has a ctypes prototype
    assert names <= set(abi.FUNCTIONS), sorted(names - set(abi.FUNCTIONS))
    nm = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in nm.splitlines() if " T " in l}
    assert names <= exported


def test_terrain_library_loads():
    from pyrit_amd import TERRAIN_PATH

    lib = C.CDLL(TERRAIN_PATH)
    for n in ("pyrgen_generate_host", "pyrgen_generate_device", "pyrgen_scratch_bytes", "pyrgen_height_host"):
        assert hasattr(lib, n)


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "pyrit.h"
#include "pyrit_hip_build.h"
#define S(T) printf(#T " size %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(PyrRay) S(PyrHit) S(PyrCamera) S(PyrGeometryData) S(PyrInstanceData) S(PyrScene) S(PyrCreateInfo)
  S(PyrVoxel) S(PyrFrameInfo) S(PyrTargets) S(PyrhipBuilt) S(PyrhipChunkBatch) S(PyrhipDag)
  S(PyrChunkKey) S(PyrWorldGenParams) S(PyrWorldInfo) S(PyrWorldStats) S(PyrWorldEdit) S(PyrDagInfo)
  S(PyrhipBuildTuning) S(PyrhipTuneReport) S(PyrMaterial) S(PyrHitEx) S(PyrStats)
  O(PyrHitEx, voxel) O(PyrHitEx, position) O(PyrHitEx, normal) O(PyrHitEx, reserved2)
  O(PyrMaterial, flags) O(PyrMaterial, texture) O(PyrMaterial, side_color) O(PyrMaterial, variation_warm)
  O(PyrDagInfo, nodes) O(PyrDagInfo, leaves) O(PyrDagInfo, attributes) O(PyrDagInfo, attribute_count)
  O(PyrWorldInfo, memory_budget) O(PyrWorldInfo, y_min) O(PyrWorldInfo, flags) O(PyrWorldInfo, generate)
  O(PyrWorldInfo, user) O(PyrWorldStats, voxel_pixels) O(PyrWorldStats, bytes) O(PyrWorldGenParams, user)
  O(PyrInstanceData, bounds_min) O(PyrInstanceData, geometry) O(PyrInstanceData, bounds_max)
  O(PyrInstanceData, mask) O(PyrInstanceData, flags)
  O(PyrScene, instance_count) O(PyrScene, frame) O(PyrScene, origin) O(PyrScene, materials)
  O(PyrScene, transparent_materials) O(PyrScene, textures) O(PyrScene, texture_count)
  O(PyrCamera, projection) O(PyrCamera, scale) O(PyrCamera, near_plane)
  O(PyrCreateInfo, cuda_context) O(PyrCreateInfo, max_instances) O(PyrCreateInfo, node_pool_bytes)
  O(PyrCreateInfo, log)
  O(PyrTargets, ray_mask) O(PyrTargets, detail_scale)
  return 0;
}
"""

# sizes the reference's ABI fixes (reference include/pyrit.h, checked there by
# tests/abi_test.zig); computed from its field lists on LP64
REFERENCE_SIZES = {"PyrRay": 32, "PyrHit": 16, "PyrCamera": 92, "PyrGeometryData": 32, "PyrInstanceData": 144,
                   "PyrScene": 224, "PyrCreateInfo": 104, "PyrVoxel": 16, "PyrFrameInfo": 32, "PyrTargets": 80,
                   "PyrChunkKey": 16, "PyrWorldGenParams": 48, "PyrWorldInfo": 88, "PyrWorldStats": 48,
                   "PyrWorldEdit": 32, "PyrDagInfo": 64, "PyrMaterial": 144, "PyrHitEx": 64,
                   "PyrStats": 72}
# field offsets the reference's world ABI fixes (reference include/pyrit.h:815-875)
REFERENCE_OFFSETS = {"PyrWorldInfo.memory_budget": 24, "PyrWorldInfo.y_min": 32, "PyrWorldInfo.flags": 68,
                     "PyrWorldInfo.generate": 72, "PyrWorldInfo.user": 80, "PyrWorldStats.voxel_pixels": 24,
                     "PyrWorldStats.bytes": 40, "PyrWorldGenParams.user": 40,
                     # reference include/pyrit.h:190-228 (field list of PyrMaterial)
                     "PyrMaterial.flags": 32, "PyrMaterial.texture": 84, "PyrMaterial.side_color": 108,
                     "PyrMaterial.variation_warm": 140}


def test_struct_layouts_match_reference_and_mirrors():
    from pyrit_amd import abi

    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write(LAYOUT_C)
        subprocess.run(["gcc", "-std=c11", f"-I{ROOT}/include", src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    vals = {}
    for line in out.splitlines():
        k, kind, v = line.rsplit(" ", 2) if " size " in line else (line.rsplit(" ", 1)[0], "off", line.rsplit(" ", 1)[1])
        vals[k] = int(v)
    for name, size in REFERENCE_SIZES.items():
        assert vals[name] == size, (name, vals[name], size)
    for name, off in REFERENCE_OFFSETS.items():
        assert vals[name] == off, (name, vals[name], off)
    mirrors = {"PyrCamera": abi.PyrCamera, "PyrGeometryData": abi.PyrGeometryData,
               "PyrInstanceData": abi.PyrInstanceData, "PyrScene": abi.PyrScene, "PyrCreateInfo": abi.PyrCreateInfo,
               "PyrFrameInfo": abi.PyrFrameInfo, "PyrTargets": abi.PyrTargets, "PyrhipBuilt": abi.PyrhipBuilt,
               "PyrhipChunkBatch": abi.PyrhipChunkBatch, "PyrhipDag": abi.PyrhipDag, "PyrChunkKey": abi.PyrChunkKey,
               "PyrWorldGenParams": abi.PyrWorldGenParams, "PyrWorldInfo": abi.PyrWorldInfo,
               "PyrWorldStats": abi.PyrWorldStats, "PyrWorldEdit": abi.PyrWorldEdit, "PyrDagInfo": abi.PyrDagInfo,
               "PyrhipBuildTuning": abi.PyrhipBuildTuning, "PyrhipTuneReport": abi.PyrhipTuneReport,
               "PyrMaterial": abi.PyrMaterial, "PyrStats": abi.PyrStats}
    for name, cls in mirrors.items():
        assert C.sizeof(cls) == vals[name], name
    for key, v in vals.items():
        if "." in key:
            sname, field = key.split(".")
            if sname in mirrors:
                assert getattr(mirrors[sname], field).offset == v, key
    assert abi.VOXEL.itemsize == vals["PyrVoxel"] and abi.RAY.itemsize == vals["PyrRay"]
    assert abi.HIT.itemsize == vals["PyrHit"]
    assert abi.HIT_EX.itemsize == vals["PyrHitEx"]
    for f in ("voxel", "position", "normal", "reserved2"):
        assert abi.HIT_EX.fields[f][1] == vals["PyrHitEx." + f], f


def test_error_strings_without_device():
    from pyrit_amd import lib

    L = lib()
    assert L.pyr_result_string(0) == b"PYR_OK"
    assert L.pyr_result_string(-3) == b"PYR_ERROR_OUT_OF_MEMORY"
    # argument validation happens before any device work
    from pyrit_amd import abi

    out = abi.PyrhipBuilt()
    assert L.pyrhip_build(2, 0, 0, 0, 0, 0, None, None, C.byref(out)) == -1
    assert b"log2_size" in L.pyr_error_message()
    assert L.pyrhip_build(5, 0, 0, 0, 0, 10, None, None, C.byref(out)) == -1


REFERENCE_INCLUDE = os.path.join(os.environ.get("PYRIT_REFERENCE", "/root/reference"), "include")
SHARED_STRUCTS = ["PyrRay", "PyrHit", "PyrCamera", "PyrGeometryData", "PyrInstanceData", "PyrScene", "PyrCreateInfo",
                  "PyrVoxel", "PyrFrameInfo", "PyrTargets", "PyrMaterial", "PyrChunkKey", "PyrWorldGenParams",
                  "PyrWorldInfo", "PyrWorldStats", "PyrWorldEdit", "PyrDagInfo", "PyrHitEx", "PyrStats"]


def _fields(name):
    from pyrit_amd import abi

    dtypes = {"PyrRay": abi.RAY, "PyrHit": abi.HIT, "PyrVoxel": abi.VOXEL, "PyrHitEx": abi.HIT_EX}
    if name in dtypes:
        return list(dtypes[name].names)
    return [f[0] for f in getattr(abi, name)._fields_]


def _layout(include_dir, structs, macros):
    lines = ["#include <stdio.h>", "#include <stddef.h>", "#include <stdint.h>", '#include "pyrit.h"', "int main(void) {"]
    for s in structs:
        lines.append(f'  printf("{s} %zu\\n", sizeof({s}));')
        for f in _fields(s):
            lines.append(f'  printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    for m in macros:
        lines.append(f'  printf("{m} %lld\\n", (long long)({m}));')
    lines.append("  return 0;\n}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        r = subprocess.run(["gcc", "-std=c11", f"-I{include_dir}", src, "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE_INCLUDE, "pyrit.h")),
                    reason="the reference's include/pyrit.h is not available here")
def test_layouts_and_constants_equal_the_reference_header():
    """Every field offset and size of the structs both headers declare, and
    every integer constant of ours that the reference header also defines,
    compiled against both headers (the check of tests/abi_test.zig:25-55,
    field by field)."""
    ref_text = open(os.path.join(REFERENCE_INCLUDE, "pyrit.h")).read()
    ours = open(os.path.join(ROOT, "include", "pyrit.h")).read()
    macros = [m for m in re.findall(r"^#define\s+(PYR_[A-Z0-9_]+)\s+\S", ours, flags=re.M)
              if re.search(r"^#define\s+" + m + r"\s", ref_text, flags=re.M) and m not in ("PYR_API",)]
    macros = [m for m in macros if not m.endswith("_H") and "(" not in m]
    assert len(macros) > 20
    a = _layout(os.path.join(ROOT, "include"), SHARED_STRUCTS, macros)
    b = _layout(REFERENCE_INCLUDE, SHARED_STRUCTS, macros)
    diff = [(x, y) for x, y in zip(a, b) if x != y]
    assert not diff, diff[:10]
    assert len(a) == len(b) and len(a) > 200
