"""MI355X-native MRG8 generator: Python mirror of the reference's C++ classes.

`mrg8_cuda` keeps the method names and argument meaning of the reference's
``mrg8`` / ``mrg8_cuda`` classes (cuda/mrg8.h:23-79, cuda/mrg8_cuda.h:20-45);
everything goes through the C ABI of libmrg8_amd.so (include/mrg8.h), whose
device work runs in hand-written gfx950 kernels.  There is no other backend:
if the library or a gfx950 device is missing, calls raise.

Formats: F64 = double(x)/p in (0,1] (original/mrg8.cpp ``rand``),
F64_01 = double(x-1)/p in [0,1) (the CUDA kernels), U32 = raw words in [1, p].
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .capi import F64, F64_01, U32, Mrg8Error  # noqa: F401

P = 2147483647


def _u32x8(state) -> "ctypes.Array":
    a = np.asarray(state, dtype=np.int64)
    if a.shape != (8,):
        raise ValueError("state must have 8 words")
    if (a < 0).any() or (a > 0xFFFFFFFF).any():
        raise ValueError("state words must be 32-bit unsigned")
    return (ctypes.c_uint32 * 8)(*[int(x) for x in a])


def _bits200(bits) -> "ctypes.Array":
    arr = (ctypes.c_short * 200)()
    for b in bits:
        if not 0 <= int(b) < 200:
            raise ValueError("bit positions must be in [0, 200)")
        arr[int(b)] = 1
    return arr


def _ptr(p) -> int:
    if hasattr(p, "data_ptr"):
        return int(p.data_ptr())  # torch tensor
    if hasattr(p, "ptr"):
        return int(p.ptr)  # mrg8_amd.hip.DeviceBuffer
    return int(p)


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "cuda_stream"):
        return ctypes.c_void_p(int(s.cuda_stream))  # torch.cuda.Stream
    if hasattr(s, "handle"):
        return ctypes.c_void_p(s.handle)
    return ctypes.c_void_p(int(s))


class mrg8_cuda:
    """One MRG8 stream bound to the current HIP device (mrg8_cuda, cuda/mrg8_cuda.h:20-45)."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        capi.call("mrg8_create", ctypes.c_uint32(seed & 0xFFFFFFFF), ctypes.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            capi.lib().mrg8_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    # -- state (mrg8::seed_init / get_state / set_state / jump_ahead) --------
    def seed_init(self, seed: int) -> None:
        capi.call("mrg8_seed_init", self._h, ctypes.c_uint32(seed & 0xFFFFFFFF))

    def get_state(self) -> np.ndarray:
        st = (ctypes.c_uint32 * 8)()
        capi.call("mrg8_get_state", self._h, st)
        return np.array(st[:], dtype=np.uint32)

    def set_state(self, state) -> None:
        capi.call("mrg8_set_state", self._h, _u32x8(state))

    def jump_ahead(self, step: int = 0, bits=None) -> None:
        """Mutating jump (original/mrg8.cpp:105-129): by `step` or by a 200-bit value given as set bit positions."""
        if bits is not None:
            capi.call("mrg8_jump_ahead_bits", self._h, _bits200(bits))
        else:
            capi.call("mrg8_jump_ahead", self._h, ctypes.c_uint64(step))

    def jump_state(self, step: int = 0, bits=None) -> np.ndarray:
        """Non-mutating jump returning the new state (cuda/mrg8.cpp:127-150)."""
        out = (ctypes.c_uint32 * 8)()
        if bits is not None:
            capi.call("mrg8_jump_state_bits", self._h, _bits200(bits), out)
        else:
            capi.call("mrg8_jump_state", self._h, ctypes.c_uint64(step), out)
        return np.array(out[:], dtype=np.uint32)

    # -- device generation ------------------------------------------------------
    def generate(self, dptr, n: int, fmt: int = F64, stream=None, block_threads: int = 0, tile: int = 0,
                 per_lane: int = 0) -> None:
        """Stream-ordered fill of n outputs at device address `dptr`; the state advances by n.

        Shape (0 = automatic, the output is the same for every shape):
        per_lane 1, 2 or 4 selects the vector kernel (block_threads a multiple
        of 64 up to 1024, tile = outputs per wave); 16 or 32 the matrix-core
        fill (int8 MFMAs over byte digits; tiles of 16 columns of 4 * per_lane
        outputs, block_threads 256, 512 or 1024, tile = outputs per wave, a
        multiple of the tile).  Automatic: the matrix-core fill for word fills
        of 2^22 outputs or more, the vector kernel otherwise."""
        if block_threads or tile or per_lane:
            capi.call("mrg8_generate_shape", self._h, ctypes.c_void_p(_ptr(dptr)), ctypes.c_uint64(n), fmt,
                      _stream(stream), block_threads, ctypes.c_uint64(tile), per_lane)
        else:
            capi.call("mrg8_generate", self._h, ctypes.c_void_p(_ptr(dptr)), ctypes.c_uint64(n), fmt,
                      _stream(stream))

    def mrg8_outer_32(self, d_ran, n: int, BS: int = 0, TNUM: int = 0) -> None:
        """cuda/mrg8_cuda.cu:414-436: blocking, [0,1) doubles.  TNUM = 0 (the
        default) takes the library's own launch shape; BS = 0 means 64, the
        reference's threads per block when only TNUM is given."""
        capi.call("mrg8_outer_32", self._h, ctypes.c_void_p(_ptr(d_ran)), n, BS, TNUM)

    def mrg8_inner(self, d_ran, n: int, TNUM: int = 64 * 1024) -> None:
        """cuda/mrg8_cuda.cu:402-412: blocking, [0,1) doubles."""
        capi.call("mrg8_inner", self._h, ctypes.c_void_p(_ptr(d_ran)), n, TNUM)

    def state_cpy_DtH(self) -> None:
        capi.call("mrg8_state_cpy_DtH", self._h)

    def synchronize(self) -> None:
        capi.call("mrg8_synchronize", self._h)

    # -- host-memory forms (class mrg8) --------------------------------------------
    def randint(self, n: int | None = None):
        if n is None:
            v = ctypes.c_uint32()
            capi.call("mrg8_randint1", self._h, ctypes.byref(v))
            return int(v.value)
        out = np.empty(n, dtype=np.uint32)
        capi.call("mrg8_randint", self._h, ctypes.c_void_p(out.ctypes.data), ctypes.c_uint64(n))
        return out

    def rand(self, n: int | None = None):
        if n is None:
            v = ctypes.c_double()
            capi.call("mrg8_rand1", self._h, ctypes.byref(v))
            return float(v.value)
        out = np.empty(n, dtype=np.float64)
        capi.call("mrg8_rand", self._h, ctypes.c_void_p(out.ctypes.data), ctypes.c_uint64(n))
        return out

    def __call__(self) -> float:
        return self.rand()

    def mrg8dnz2(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        capi.call("mrg8_mrg8dnz2", self._h, ctypes.c_void_p(out.ctypes.data), ctypes.c_uint64(n))
        return out


# The host class of the reference has the same surface minus the device launchers.
mrg8 = mrg8_cuda


def seed_state(seed: int) -> np.ndarray:
    """State after seeding (original/mrg8.cpp:180-194)."""
    st = (ctypes.c_uint32 * 8)()
    capi.call("mrg8_seed_state", ctypes.c_uint32(seed & 0xFFFFFFFF), st)
    return np.array(st[:], dtype=np.uint32)


def advance_state(state, step: int = 0, bits=None) -> np.ndarray:
    """`state` advanced by `step` positions (or by a 200-bit value given as set bits)."""
    out = (ctypes.c_uint32 * 8)()
    if bits is not None:
        capi.call("mrg8_advance_state_bits", _u32x8(state), _bits200(bits), out)
    else:
        capi.call("mrg8_advance_state", _u32x8(state), ctypes.c_uint64(step), out)
    return np.array(out[:], dtype=np.uint32)


def tune_fill(fmt: int, n: int, cache_path: str | None = None, device: int = 0):
    """Autotunes the single-stream fill shape for (fmt, size class of n) on `device`."""
    sh = capi.FillShape()
    capi.call("mrg8_tune_fill", device, fmt, ctypes.c_uint64(n), cache_path.encode() if cache_path else None,
              ctypes.byref(sh))
    return {"block_threads": sh.block_threads, "outputs_per_lane": sh.outputs_per_lane,
            "outputs_per_wave": sh.outputs_per_wave}


def jump_matrix() -> np.ndarray:
    """A^(2^j) mod p, j = 0..246, column-major per matrix (jump_matrix.txt layout)."""
    t = np.zeros(247 * 64, dtype=np.uint32)
    capi.call("mrg8_jump_matrix", t.ctypes.data_as(capi.u32p), ctypes.c_size_t(t.size))
    return t


def write_jump_matrix(path: str) -> None:
    capi.call("mrg8_write_jump_matrix", path.encode())


def read_jump_matrix(path: str) -> np.ndarray:
    """Parses a jump_matrix.txt (original/mrg8.cpp:88-102) into the layout of jump_matrix()."""
    t = np.zeros(247 * 64, dtype=np.uint32)
    capi.call("mrg8_read_jump_matrix", path.encode(), t.ctypes.data_as(capi.u32p), ctypes.c_size_t(t.size))
    return t


def check_jump_matrix(path: str) -> None:
    """Raises Mrg8Error unless `path` holds exactly the computed table."""
    capi.call("mrg8_check_jump_matrix", path.encode())


def advance_states(d_in, d_offsets, words: int, d_out, n: int, stream=None, device: int = 0) -> None:
    """Device-side bulk jump: n states (8 uint32 words each, device memory)
    advanced by little-endian offsets of `words` uint64 words each (up to 256
    bits); stream-ordered.  See mrg8_advance_states in include/mrg8.h."""
    capi.call("mrg8_advance_states", device, ctypes.c_void_p(_ptr(d_in)), ctypes.c_void_p(_ptr(d_offsets)),
              ctypes.c_uint32(words), ctypes.c_void_p(_ptr(d_out)), ctypes.c_uint64(n), _stream(stream))


class Batch:
    """Many stream segments generated by one launch (a plain kernel launch,
    or with use_graph(True) a replay of a captured HIP graph)."""

    def __init__(self, max_streams: int, max_tiles: int = 1 << 21, device: int = 0):
        h = ctypes.c_void_p()
        capi.call("mrg8_batch_create", device, ctypes.c_uint32(max_streams), ctypes.c_uint64(max_tiles),
                  ctypes.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            capi.lib().mrg8_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    @property
    def config(self):
        c = capi.BatchConfig()
        capi.call("mrg8_batch_get_config", self._h, ctypes.byref(c))
        return {"block_threads": c.block_threads, "blocks_per_cu": c.blocks_per_cu, "tile_outputs": c.tile_outputs}

    def set_config(self, block_threads: int, blocks_per_cu: int, tile_outputs: int) -> None:
        """blocks_per_cu 0: one tile per wave; < 0: a persistent grid of CUs x
        -blocks_per_cu blocks taking the tiles in address order; > 0: a
        persistent grid striding over the tiles.  The outputs are the same."""
        c = capi.BatchConfig(block_threads, blocks_per_cu, tile_outputs)
        capi.call("mrg8_batch_set_config", self._h, ctypes.byref(c))

    def set(self, segments) -> None:
        """segments: iterable of (state[8], offset, count, device_ptr, fmt)."""
        segs = list(segments)
        arr = (capi.StreamDesc * max(1, len(segs)))()
        for i, (st, off, cnt, ptr, fmt) in enumerate(segs):
            d = arr[i]
            d.state = _u32x8(st)
            d.offset = int(off)
            d.count = int(cnt)
            d.out = _ptr(ptr)
            d.format = int(fmt)
            d.reserved = 0
        capi.call("mrg8_batch_set", self._h, arr, ctypes.c_uint32(len(segs)))

    def run(self, stream=None) -> None:
        capi.call("mrg8_batch_run", self._h, _stream(stream))

    def use_graph(self, enable: bool = True) -> None:
        capi.call("mrg8_batch_use_graph", self._h, 1 if enable else 0)

    def autotune(self, stream=None, cache_path: str | None = None):
        c = capi.BatchConfig()
        capi.call("mrg8_batch_autotune", self._h, _stream(stream), cache_path.encode() if cache_path else None,
                  ctypes.byref(c))
        return {"block_threads": c.block_threads, "blocks_per_cu": c.blocks_per_cu, "tile_outputs": c.tile_outputs}


class Server:
    """Serving loop with device-resident stream states (see include/mrg8.h)."""

    def __init__(self, num_streams: int, max_output_bytes: int, base_seed: int = 1, max_inflight: int = 4096,
                 step_outputs: int = 0, tile_outputs: int = 0, block_threads: int = 0, blocks_per_cu: int = 0,
                 device: int = 0, steps_in_flight: int = 0, age_limit_us: int = 0):
        cfg = capi.ServerConfig(num_streams, max_inflight, max_output_bytes, step_outputs, tile_outputs,
                                block_threads, blocks_per_cu, steps_in_flight, age_limit_us)
        h = ctypes.c_void_p()
        capi.call("mrg8_server_create", device, ctypes.byref(cfg), ctypes.c_uint32(base_seed), ctypes.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            capi.lib().mrg8_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_id: int, state, position: int = 0) -> None:
        capi.call("mrg8_server_set_stream", self._h, ctypes.c_uint32(stream_id), _u32x8(state),
                  ctypes.c_uint64(position))

    def set_streams(self, states, positions=None, first: int = 0) -> None:
        """Sets streams first .. first+len(states)-1 in one upload."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.uint32).reshape(-1, 8))
        pos = None
        if positions is not None:
            pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint64))
            assert pos.size == st.shape[0]
        capi.call("mrg8_server_set_streams", self._h, ctypes.c_uint32(first), ctypes.c_uint32(st.shape[0]),
                  st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                  ctypes.c_void_p(pos.ctypes.data) if pos is not None else None)

    def get_stream(self, stream_id: int):
        st = (ctypes.c_uint32 * 8)()
        pos = ctypes.c_uint64()
        capi.call("mrg8_server_get_stream", self._h, ctypes.c_uint32(stream_id), st, ctypes.byref(pos))
        return np.array(st[:], dtype=np.uint32), int(pos.value)

    def submit(self, stream_id: int, count: int, fmt: int = F64) -> int:
        rid = ctypes.c_uint64()
        capi.call("mrg8_server_submit", self._h, ctypes.c_uint32(stream_id), ctypes.c_uint64(count), fmt,
                  ctypes.byref(rid))
        return int(rid.value)

    def progress(self, max_done: int = 256, wait_ms: int = 0):
        """Returns a list of completion dicts (request_id, stream_id, start, count, out, times)."""
        buf = (capi.Completion * max(1, max_done))()
        nd = ctypes.c_uint32()
        capi.call("mrg8_server_progress", self._h, buf, ctypes.c_uint32(max_done), ctypes.byref(nd), wait_ms)
        return [{"request_id": c.request_id, "stream_id": c.stream_id, "status": c.status, "start": c.start,
                 "count": c.count, "out": c.out, "submit_time_s": c.submit_time_s,
                 "complete_time_s": c.complete_time_s} for c in buf[:nd.value]]

    def release(self, request_id: int) -> None:
        capi.call("mrg8_server_release", self._h, ctypes.c_uint64(request_id))

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        capi.call("mrg8_server_stats", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("queued", "inflight", "completed", "bytes_inflight"), [x.value for x in v]))

    def counters(self):
        c = capi.ServerCounters()
        capi.call("mrg8_server_counters_get", self._h, ctypes.byref(c))
        return {k: getattr(c, k) for k, _ in capi.ServerCounters._fields_}

    def replay(self, arrivals, streams, counts, fmt: int = F64, want_latencies: bool = False):
        """Replays a trace (arrival seconds, stream ids, sizes); returns stats (and latencies)."""
        n = len(arrivals)
        tr = (capi.TraceEntry * max(1, n))()
        for i in range(n):
            tr[i].arrival_s = float(arrivals[i])
            tr[i].stream_id = int(streams[i])
            tr[i].format = fmt
            tr[i].count = int(counts[i])
        lat = np.zeros(max(1, n), dtype=np.float64)
        st = capi.TraceStats()
        capi.call("mrg8_server_replay", self._h, tr, ctypes.c_uint64(n),
                  lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if want_latencies else None,
                  ctypes.byref(st))
        out = {k: getattr(st, k) for k, _ in capi.TraceStats._fields_}
        return (out, lat[:n]) if want_latencies else out
