"""Minimal ctypes binding to the HIP runtime: device buffers, copies, streams
and events for driving libmrg8_amd from Python without importing torch.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def _load():
    global _hip
    if _hip is not None:
        return _hip
    cands = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
             "/opt/rocm/lib/libamdhip64.so", ctypes.util.find_library("amdhip64")]
    last = None
    for c in cands:
        if not c:
            continue
        try:
            h = ctypes.CDLL(c, mode=ctypes.RTLD_GLOBAL)
            break
        except OSError as e:  # pragma: no cover - environment dependent
            last = e
    else:
        raise RuntimeError(f"libamdhip64.so not loadable: {last}")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    h.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    h.hipMemsetD8.argtypes = [vp, ctypes.c_ubyte, sz]
    h.hipMemsetAsync.argtypes = [vp, ctypes.c_int, sz, vp]
    h.hipDeviceSynchronize.argtypes = []
    h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipStreamDestroy.argtypes = [vp]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    h.hipEventDestroy.argtypes = [vp]
    h.hipEventRecord.argtypes = [vp, vp]
    h.hipEventSynchronize.argtypes = [vp]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    h.hipGetErrorString.argtypes = [ctypes.c_int]
    h.hipGetErrorString.restype = ctypes.c_char_p
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
    _hip = h
    return h


class HipError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _load().hipGetErrorString(rc)
        raise HipError(f"{what} failed: {rc} {msg.decode() if msg else ''}")


def device_count() -> int:
    try:
        h = _load()
    except RuntimeError:
        return 0
    n = ctypes.c_int(0)
    if h.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def synchronize() -> None:
    check(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


def memset_async(ptr: int, value: int, nbytes: int, stream=None) -> None:
    """Stream-ordered byte fill (the runtime's own fill kernel)."""
    sp = ctypes.c_void_p(stream.handle if isinstance(stream, Stream) else stream)
    check(_load().hipMemsetAsync(ctypes.c_void_p(int(ptr)), int(value), int(nbytes), sp), "hipMemsetAsync")


def mem_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(_load().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """An owning device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(_load().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            _load().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.free()
        except Exception:
            pass

    def __int__(self):
        return int(self.ptr)

    def copy_to_host(self, dtype, count: int, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        nb = out.nbytes
        assert offset_bytes + nb <= self.nbytes
        check(_load().hipMemcpy(out.ctypes.data, ctypes.c_void_p(self.ptr + offset_bytes), nb, _D2H), "hipMemcpy D2H")
        return out

    def copy_from_host(self, arr: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        assert offset_bytes + a.nbytes <= self.nbytes
        check(_load().hipMemcpy(ctypes.c_void_p(self.ptr + offset_bytes), a.ctypes.data, a.nbytes, _H2D),
              "hipMemcpy H2D")

    def fill_byte(self, value: int) -> None:
        check(_load().hipMemsetD8(ctypes.c_void_p(self.ptr), value & 0xFF, self.nbytes), "hipMemsetD8")


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(_load().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        check(_load().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            _load().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(_load().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None) -> None:
        s = stream.handle if isinstance(stream, Stream) else stream
        check(_load().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(s)), "hipEventRecord")

    def synchronize(self) -> None:
        check(_load().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(_load().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle)),
              "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self) -> None:
        if self.handle:
            _load().hipEventDestroy(ctypes.c_void_p(self.handle))
            self.handle = None
