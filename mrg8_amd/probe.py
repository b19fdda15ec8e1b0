"""ctypes binding to libmrg8_probe.so: the benchmark's write-bandwidth probe
(the counterpart of the reference's cuda/stream.cu) and shader-clock probe
(include/mrg8_probe.h).  Measurement only; not part of the generator."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrg8_probe.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with `python -m mrg8_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.mrg8_probe_write.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_void_p]
        L.mrg8_probe_sclk.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mrg8_probe_mad64.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]
        L.mrg8_probe_xcc.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        _lib = L
    return _lib


def _s(stream):
    return ctypes.c_void_p(stream.handle if hasattr(stream, "handle") else stream)


def xcc_ids(blocks: int = 4096) -> np.ndarray:
    """The XCD each of `blocks` one-wave blocks ran on (block order)."""
    from . import hip

    buf = hip.DeviceBuffer(4 * blocks)
    try:
        buf.fill_byte(0xFF)  # a block that wrote nothing leaves 0xffffffff
        rc = lib().mrg8_probe_xcc(ctypes.c_void_p(buf.ptr), blocks, None)
        if rc != 0:
            raise RuntimeError(f"mrg8_probe_xcc failed ({rc})")
        hip.synchronize()
        return buf.copy_to_host(np.uint32, blocks)
    finally:
        buf.free()


def write(dptr: int, n: int, mode: int = 1, tile: int = 0, stream=None) -> None:
    """d[i] = i for n doubles at device address dptr (mode 0: one element per
    thread as cuda/stream.cu; mode 1: the double fill's 1 KiB non-temporal
    wave stores, 512-thread blocks in XCD runs; mode 2: the same runs in
    1024-thread blocks in dispatch order)."""
    rc = lib().mrg8_probe_write(ctypes.c_void_p(int(dptr)), n, mode, tile, _s(stream))
    if rc != 0:
        raise RuntimeError(f"mrg8_probe_write failed ({rc})")


def mad64_rate(stream, blocks: int = 4096, iters: int = 2048, reps: int = 5) -> float:
    """Sustained 32x32->64 multiply-adds per second over the whole GPU
    (eight independent v_mad_i64_i32 chains per lane)."""
    from . import hip

    out = hip.DeviceBuffer(blocks * 256 * 8)
    try:
        launch = lambda: lib().mrg8_probe_mad64(ctypes.c_void_p(out.ptr), blocks, iters, _s(stream))  # noqa: E731
        if launch() != 0:
            raise RuntimeError("mrg8_probe_mad64 failed")
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        e0.destroy()
        e1.destroy()
    finally:
        out.free()
    return blocks * 256 * 8 * iters / (ms * 1e-3)


class ClockProbe:
    """Shader clock (MHz) sampled by a one-wave kernel queued on a stream."""

    def __init__(self, capacity: int):
        from . import hip

        self.capacity = capacity
        self.buf = hip.DeviceBuffer(16 * capacity)
        self.n = 0

    def sample(self, stream) -> None:
        if self.n >= self.capacity:
            raise IndexError("clock probe buffer full")
        rc = lib().mrg8_probe_sclk(ctypes.c_void_p(self.buf.ptr + 16 * self.n), _s(stream))
        if rc != 0:
            raise RuntimeError(f"mrg8_probe_sclk failed ({rc})")
        self.n += 1

    def mhz(self) -> list:
        v = self.buf.copy_to_host(np.uint64, 2 * self.n).reshape(-1, 2)
        return [round(100.0 * float(t) / float(r), 1) if r else 0.0 for t, r in v]

    def free(self) -> None:
        self.buf.free()
