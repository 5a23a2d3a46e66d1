"""Minimal HIP runtime plumbing for Python callers: device buffers, copies and
stream sync through libamdhip64 (no torch needed)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def _ok(code: int, what: str) -> None:
    if code != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(code).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    try:
        if hip().hipGetDeviceCount(C.byref(n)) != 0:
            return 0
    except RuntimeError:
        return 0
    return n.value


def mem_info() -> tuple[int, int]:
    free, total = C.c_size_t(0), C.c_size_t(0)
    _ok(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class Buffer:
    """An owned device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 16)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray) -> "Buffer":
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            _ok(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "upload")
        return self

    def zero(self) -> "Buffer":
        _ok(hip().hipMemset(C.c_void_p(self.ptr), 0, self.nbytes), "memset")
        return self

    def download(self, dtype, count: int | None = None) -> np.ndarray:
        return download(self.ptr, dtype, count if count is not None else self.nbytes // np.dtype(dtype).itemsize)


def to_device(arr: np.ndarray) -> Buffer:
    a = np.ascontiguousarray(arr)
    return Buffer(a.nbytes).upload(a)


def download(ptr: int, dtype, count: int) -> np.ndarray:
    out = np.empty(int(count), dtype=dtype)
    if out.nbytes:
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2), "download")
    return out


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def sync(self) -> None:
        _ok(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            hip().hipStreamDestroy(C.c_void_p(self.handle))
            self.handle = 0


class Events:
    """HIP events recorded on a stream; elapsed() in ms."""

    def __init__(self, stream: Stream | None = None):
        self.stream = stream.handle if stream else None
        a, b = C.c_void_p(), C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(a)), "event")
        _ok(hip().hipEventCreate(C.byref(b)), "event")
        self.a, self.b = a.value, b.value

    def start(self) -> None:
        _ok(hip().hipEventRecord(C.c_void_p(self.a), C.c_void_p(self.stream)), "record")

    def stop(self) -> float:
        _ok(hip().hipEventRecord(C.c_void_p(self.b), C.c_void_p(self.stream)), "record")
        _ok(hip().hipEventSynchronize(C.c_void_p(self.b)), "event sync")
        ms = C.c_float(0)
        _ok(hip().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.a), C.c_void_p(self.b)), "elapsed")
        return ms.value

    def destroy(self) -> None:
        hip().hipEventDestroy(C.c_void_p(self.a))
        hip().hipEventDestroy(C.c_void_p(self.b))


def gpu_available() -> bool:
    if os.environ.get("PYRIT_FORCE_NO_GPU"):
        return False
    return device_count() > 0
