"""rocRAND generators timed beside MRG8 on the same GPU, for context.

The counterpart of the reference's cuRAND drivers
(cuda/tests/test_curand_{xorwow,mrg32k3a,mtgp32,philox4_32_10,mt19937}.cu:
seed 13579 set before every run, N doubles from the generator's uniform-double
call timed with events, ITER runs, the first dropped, mean of the rest).
Two figures per generator:

* ``reseed_ms``: the drivers' own method -- the seed is set before every run,
  so a run includes the generator's state set-up;
* ``steady_ms``: back-to-back calls on a seeded generator (what a long-running
  consumer sees), comparable to MRG8's per-call time.

Different generators produce different sequences: these are rates, not
parity checks.
"""
from __future__ import annotations

import ctypes
import os

from . import hip

GENERATORS = {
    "xorwow": 401,
    "mrg32k3a": 402,
    "mtgp32": 403,
    "philox4_32_10": 404,
    "mt19937": 407,
}

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocrand.so")
    h = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    h.rocrand_create_generator.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    h.rocrand_destroy_generator.argtypes = [vp]
    h.rocrand_set_seed.argtypes = [vp, ctypes.c_ulonglong]
    h.rocrand_set_stream.argtypes = [vp, vp]
    h.rocrand_generate_uniform_double.argtypes = [vp, vp, ctypes.c_size_t]
    _lib = h
    return h


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with rocrand status {rc}")


def time_generator(name: str, buf: hip.DeviceBuffer, n: int, iters: int = 10, seed: int = 13579,
                   stream: hip.Stream | None = None) -> dict:
    """Times one rocRAND generator filling ``n`` doubles of ``buf``."""
    lib = _load()
    s = stream or hip.Stream()
    gen = ctypes.c_void_p()
    _check(lib.rocrand_create_generator(ctypes.byref(gen), GENERATORS[name]), f"create {name}")
    try:
        _check(lib.rocrand_set_stream(gen, ctypes.c_void_p(s.handle)), "set_stream")
        e0, e1 = hip.Event(), hip.Event()
        dst = ctypes.c_void_p(buf.ptr)

        reseed = []
        for _ in range(iters):
            _check(lib.rocrand_set_seed(gen, seed), "set_seed")
            s.synchronize()
            e0.record(s)
            _check(lib.rocrand_generate_uniform_double(gen, dst, n), f"generate {name}")
            e1.record(s)
            e1.synchronize()
            reseed.append(e0.elapsed_ms(e1))
        reseed_ms = sum(reseed[1:]) / (len(reseed) - 1)

        _check(lib.rocrand_generate_uniform_double(gen, dst, n), f"generate {name}")
        s.synchronize()
        e0.record(s)
        for _ in range(iters):
            _check(lib.rocrand_generate_uniform_double(gen, dst, n), f"generate {name}")
        e1.record(s)
        e1.synchronize()
        steady_ms = e0.elapsed_ms(e1) / iters
    finally:
        lib.rocrand_destroy_generator(gen)
        if stream is None:
            s.destroy()
    return {"reseed_ms": reseed_ms, "steady_ms": steady_ms,
            "reseed_gsps": n / (reseed_ms * 1e-3) / 1e9, "steady_gsps": n / (steady_ms * 1e-3) / 1e9}


def rocrand_rates(n: int, iters: int = 10, names=None) -> dict:
    """All (or the named) generators, each filling ``n`` doubles."""
    buf = hip.DeviceBuffer(n * 8)
    out = {}
    try:
        for name in names or GENERATORS:
            try:
                out[name] = time_generator(name, buf, n, iters)
            except (OSError, RuntimeError) as e:
                out[name] = {"error": str(e)}
    finally:
        buf.free()
    return out
