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
