"""MRG8 oracle: ctypes bindings to the plain-C restatement in this directory.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker (or the timed CPU
baseline), never as the thing measured or shipped.  The product package
``mrg8_amd`` never imports this module.

The C sources cite the reference file:line each function restates
(original/mrg8.cpp, cuda/mrg8.cpp, knl/mrg8.h, knl/mrg8_vec.h).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrg8_oracle.so")
P = 2147483647
NJUMP = 247
FNV_INIT = 0xCBF29CE484222325
FMT_F64 = 0      # double(x) * (1/p), range (0, 1]   (original/mrg8.cpp:233-236)
FMT_F64_01 = 1   # double(x-1) * (1/p), range [0, 1) (cuda/mrg8_cuda.cu:363)
FMT_U32 = 2      # raw word in [1, p]                (original/mrg8.cpp:219)

_lib = None


def build() -> None:
    """Compile the oracle library (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.mrg8o_bigdot.argtypes = [u32p, u32p]
        L.mrg8o_bigdot.restype = ctypes.c_uint32
        L.mrg8o_seed.argtypes = [ctypes.c_uint32, u32p]
        for f in (L.mrg8o_randint, L.mrg8o_rand, L.mrg8o_rand01, L.mrg8o_mrg8dnz2):
            f.argtypes = [u32p, ctypes.c_void_p, ctypes.c_uint64]
            f.restype = None
        L.mrg8o_jump_table.argtypes = [u32p]
        L.mrg8o_write_jump_table.argtypes = [u32p, ctypes.c_char_p]
        L.mrg8o_write_jump_table.restype = ctypes.c_int
        L.mrg8o_jump_bits.argtypes = [u32p, ctypes.POINTER(ctypes.c_short), u32p]
        L.mrg8o_jump.argtypes = [u32p, ctypes.c_uint64, u32p]
        L.mrg8o_jump_vec.argtypes = [u32p, ctypes.c_uint64, u32p, u32p]
        L.mrg8o_fnv1a.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t]
        L.mrg8o_fnv1a.restype = ctypes.c_uint64
        L.mrg8o_hash_stream.argtypes = [u32p, ctypes.c_uint64, ctypes.c_int]
        L.mrg8o_hash_stream.restype = ctypes.c_uint64
        L.mrg8o_wsum.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        L.mrg8o_wsum.restype = ctypes.c_uint64
        L.mrg8o_stream_wsums.argtypes = [u32p, u32p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.c_void_p]
        L.mrg8o_stream_wsums.restype = None
        L.mrg8o_stream_windows.argtypes = [u32p, u32p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_int, ctypes.c_void_p]
        L.mrg8o_stream_windows.restype = None
        L.mrg8o_vec_outer_tp.argtypes = [u32p, u32p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.mrg8o_vec_outer_tp.restype = ctypes.c_int
        L.mrg8o_vec_has_avx512.restype = ctypes.c_int
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def coeffs() -> np.ndarray:
    return np.array([1089656042, 1906537547, 1764115693, 1304127872,
                     189748160, 1984088114, 626062218, 1927846343], dtype=np.uint32)


def bigdot(x, y) -> int:
    x, y = _u32(x), _u32(y)
    return int(lib().mrg8o_bigdot(_ptr(x), _ptr(y)))


def seed_state(seed: int) -> np.ndarray:
    st = np.zeros(8, dtype=np.uint32)
    lib().mrg8o_seed(ctypes.c_uint32(seed & 0xFFFFFFFF), _ptr(st))
    return st


def generate(state, n: int, fmt: int = FMT_F64):
    """Return (outputs, new_state) for n outputs from `state` (newest word first)."""
    st = _u32(state).copy()
    if fmt == FMT_U32:
        out = np.empty(n, dtype=np.uint32)
        lib().mrg8o_randint(_ptr(st), out.ctypes.data, n)
    else:
        out = np.empty(n, dtype=np.float64)
        f = lib().mrg8o_rand if fmt == FMT_F64 else lib().mrg8o_rand01
        f(_ptr(st), out.ctypes.data, n)
    return out, st


def mrg8dnz2(state, n: int):
    st = _u32(state).copy()
    out = np.empty(n, dtype=np.float64)
    lib().mrg8o_mrg8dnz2(_ptr(st), out.ctypes.data, n)
    return out, st


_table = None


def jump_table() -> np.ndarray:
    global _table
    if _table is None:
        t = np.zeros(64 * NJUMP, dtype=np.uint32)
        lib().mrg8o_jump_table(_ptr(t))
        _table = t
    return _table


def jump(state, step: int) -> np.ndarray:
    """Reference jump_ahead(u64): J = prod of table matrices, then J * state."""
    st = _u32(state).copy()
    lib().mrg8o_jump(_ptr(jump_table()), ctypes.c_uint64(step), _ptr(st))
    return st


def jump_bits(state, bits) -> np.ndarray:
    """200-bit jump; `bits` is an iterable of set bit positions (< 200)."""
    b = (ctypes.c_short * 200)()
    for i in bits:
        b[i] = 1
    st = _u32(state).copy()
    lib().mrg8o_jump_bits(_ptr(jump_table()), b, _ptr(st))
    return st


def jump_vec(state, step: int) -> np.ndarray:
    """knl-style matrix-vector jump chain."""
    st = _u32(state).copy()
    out = np.zeros(8, dtype=np.uint32)
    lib().mrg8o_jump_vec(_ptr(jump_table()), ctypes.c_uint64(step), _ptr(st), _ptr(out))
    return out


def fnv1a(buf: np.ndarray, h: int = FNV_INIT) -> int:
    b = np.ascontiguousarray(buf)
    return int(lib().mrg8o_fnv1a(ctypes.c_uint64(h), b.ctypes.data, b.nbytes))


def hash_stream(state, n: int, fmt: int):
    """(fnv1a64 of n outputs in format fmt, new state) without holding the outputs."""
    st = _u32(state).copy()
    h = int(lib().mrg8o_hash_stream(_ptr(st), n, fmt))
    return h, st


def vec_outer_tp(state, n: int, nthreads: int = 0, out: np.ndarray | None = None):
    """CPU baseline (knl vector outer product, OpenMP).  Returns (out, new_state, threads)."""
    st = _u32(state).copy()
    if out is None:
        out = np.empty(n, dtype=np.float64)
    assert out.dtype == np.float64 and out.size >= n and out.flags.c_contiguous
    used = lib().mrg8o_vec_outer_tp(_ptr(jump_table()), _ptr(st), out.ctypes.data, n, nthreads)
    return out, st, used


def vec_has_avx512() -> bool:
    return bool(lib().mrg8o_vec_has_avx512())


def wsum(buf: np.ndarray) -> int:
    """sum_i w_i (2i+1) mod 2^64 over the outputs (u32 words or f64 bit patterns)."""
    b = np.ascontiguousarray(buf)
    return int(lib().mrg8o_wsum(b.ctypes.data, b.size, b.itemsize))


def stream_wsums(states, offsets, counts, formats) -> np.ndarray:
    """wsum of each stream's outputs: state k jumped by offsets[k], counts[k] outputs."""
    st = np.ascontiguousarray(np.asarray(states, dtype=np.uint32).reshape(-1, 8))
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    cnt = np.ascontiguousarray(np.asarray(counts, dtype=np.uint64))
    fmt = np.ascontiguousarray(np.asarray(formats, dtype=np.int32))
    out = np.zeros(len(cnt), dtype=np.uint64)
    lib().mrg8o_stream_wsums(_ptr(jump_table()), _ptr(st), off.ctypes.data, cnt.ctypes.data, fmt.ctypes.data,
                             len(cnt), out.ctypes.data)
    return out


def stream_windows(states, offsets, m: int, fmt: int) -> np.ndarray:
    """(n, m) outputs: state k jumped by offsets[k], then its next m outputs in format fmt."""
    st = np.ascontiguousarray(np.asarray(states, dtype=np.uint32).reshape(-1, 8))
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    assert off.size == st.shape[0]
    out = np.empty((st.shape[0], m), dtype=np.uint32 if fmt == FMT_U32 else np.float64)
    lib().mrg8o_stream_windows(_ptr(jump_table()), _ptr(st), off.ctypes.data, st.shape[0], m, fmt,
                               out.ctypes.data)
    return out
