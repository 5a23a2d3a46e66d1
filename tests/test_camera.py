"""Camera helpers of the product on the CPU (no GPU needed): look-at and
perspective (src/root.zig:458-481) equal the oracle's restatement byte for
byte; orthographic (src/root.zig:483-489) fills projection, size and scale
as stated there and leaves the pose, shift, jitter and planes alone."""
import ctypes as C

import numpy as np

import oracle_py as O


def test_look_at_and_perspective_equal_the_oracle():
    from pyrit_amd import abi, lib

    L = lib()
    for eye, tgt, up, fov, w, h, near in [((-32, 96, -32), (32, 32, 32), (0, 1, 0), np.pi / 3, 256, 256, 0.0),
                                          ((5, -3, 200), (0, 0, 0), (0, 0, 1), 0.4, 1920, 1080, 0.25)]:
        c = abi.PyrCamera()
        L.pyr_camera_look_at(C.byref(c), (C.c_float * 3)(*eye), (C.c_float * 3)(*tgt), (C.c_float * 3)(*up))
        L.pyr_camera_perspective(C.byref(c), fov, w, h, near)
        assert bytes(c) == bytes(O.camera(eye, tgt, up, fov, w, h, near))


def test_orthographic_fields():
    from pyrit_amd import abi, lib

    L = lib()
    c = abi.PyrCamera()
    L.pyr_camera_look_at(C.byref(c), (C.c_float * 3)(1, 2, 3), (C.c_float * 3)(4, 0, 9), (C.c_float * 3)(0, 1, 0))
    c.shift[:] = [0.125, -0.25]
    c.jitter[:] = [0.5, 0.5]
    c.near_plane, c.far_plane = 2.0, 90.0
    pose = list(c.view_to_world)
    L.pyr_camera_orthographic(C.byref(c), 37.5, 300, 200)
    assert c.projection == abi.PYR_PROJECTION_ORTHOGRAPHIC == 1
    assert (c.width, c.height) == (300, 200)
    assert c.scale[0] == np.float32(np.float32(37.5) * np.float32(300.0) / np.float32(200.0)) and c.scale[1] == 37.5
    assert list(c.view_to_world) == pose and list(c.shift) == [0.125, -0.25] and list(c.jitter) == [0.5, 0.5]
    assert (c.near_plane, c.far_plane) == (2.0, 90.0)
    # back to a perspective camera
    L.pyr_camera_perspective(C.byref(c), 1.0, 300, 200, 0.0)
    assert c.projection == abi.PYR_PROJECTION_PERSPECTIVE == 0


def _radical_inverse(i, b):
    """the base-b digits of i mirrored behind the point, exactly (fractions)"""
    from fractions import Fraction

    r, f = Fraction(0), Fraction(1)
    while i:
        f /= b
        r += f * (i % b)
        i //= b
    return r


def test_jitter_sequences():
    """pyr_jitter_halton (root.zig:427-443): the (2, 3) radical inverses of
    frame % 1024 + 1, minus 1/2, within float rounding of the exact values,
    repeating after 1024 frames; pyr_jitter_ordered (root.zig:405-425): the
    16 frames visit the centres of a 4x4 grid once each, the first four
    frames take rows 0, 2, 1, 3, odd rows run backwards, period 16."""
    from pyrit_amd import lib

    L = lib()
    out = (C.c_float * 2)()

    def jit(fn, frame):
        fn(frame, out)
        return float(out[0]), float(out[1])

    assert jit(L.pyr_jitter_halton, 0) == (0.0, float(np.float32(1) / np.float32(3) - np.float32(0.5)))
    for frame in list(range(40)) + [1023, 1024, 5000]:
        i = frame % 1024 + 1
        x, y = jit(L.pyr_jitter_halton, frame)
        assert abs(x - float(_radical_inverse(i, 2) - 0.5)) < 1e-6
        assert abs(y - float(_radical_inverse(i, 3) - 0.5)) < 1e-6
        assert -0.5 <= x < 0.5 and -0.5 <= y < 0.5
    assert jit(L.pyr_jitter_halton, 7) == jit(L.pyr_jitter_halton, 7 + 1024)
    cells = [jit(L.pyr_jitter_ordered, f) for f in range(16)]
    centres = {((c + 0.5) / 4 - 0.5, (r + 0.5) / 4 - 0.5) for c in range(4) for r in range(4)}
    assert set(cells) == centres
    rows = [int(round((y + 0.5) * 4 - 0.5)) for _, y in cells]
    cols = [int(round((x + 0.5) * 4 - 0.5)) for x, _ in cells]
    assert rows[:4] == [0, 2, 1, 3] and rows[4:8] == [0, 2, 1, 3]
    assert [cols[k] for k in range(0, 16, 4)] == [0, 1, 2, 3]  # row 0 forwards
    assert [cols[k] for k in range(2, 16, 4)] == [3, 2, 1, 0]  # row 1 backwards
    assert all(jit(L.pyr_jitter_ordered, f) == cells[f % 16] for f in range(16, 64))
