"""ctypes declarations for libmrg8_amd.so (the C ABI in include/mrg8.h).

The library is required: importing this module raises if it is missing, so
no caller can silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrg8_amd.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NODEVICE, ERR_CAPACITY, ERR_STATE, ERR_NOTFOUND, ERR_IO = -1, -2, -3, -4, -5, -6, -7, -8
ERR_BLOCKED = -9
F64, F64_01, U32 = 0, 1, 2


class StreamDesc(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32 * 8), ("offset", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("out", ctypes.c_void_p), ("format", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("block_threads", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("tile_outputs", ctypes.c_uint64)]


class FillShape(ctypes.Structure):
    _fields_ = [("block_threads", ctypes.c_int32), ("outputs_per_lane", ctypes.c_int32),
                ("outputs_per_wave", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("num_streams", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32),
                ("max_output_bytes", ctypes.c_uint64), ("step_outputs", ctypes.c_uint64),
                ("tile_outputs", ctypes.c_uint64), ("block_threads", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("steps_in_flight", ctypes.c_uint32),
                ("age_limit_us", ctypes.c_uint32)]


class Completion(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_uint64), ("stream_id", ctypes.c_uint32), ("status", ctypes.c_int32),
                ("start", ctypes.c_uint64), ("count", ctypes.c_uint64), ("out", ctypes.c_void_p),
                ("submit_time_s", ctypes.c_double), ("complete_time_s", ctypes.c_double)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("step_outputs", ctypes.c_uint64), ("step_segments", ctypes.c_uint64),
                ("host_launch_s", ctypes.c_double), ("max_queue_s", ctypes.c_double)]


class TraceEntry(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("stream_id", ctypes.c_uint32), ("format", ctypes.c_int32),
                ("count", ctypes.c_uint64)]


class TraceStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("outputs", ctypes.c_uint64), ("wall_s", ctypes.c_double),
                ("p50_s", ctypes.c_double), ("p90_s", ctypes.c_double), ("p99_s", ctypes.c_double),
                ("max_s", ctypes.c_double), ("mean_s", ctypes.c_double), ("max_host_gap_s", ctypes.c_double),
                ("max_host_gap_at_s", ctypes.c_double)]


u32p = ctypes.POINTER(ctypes.c_uint32)
vp = ctypes.c_void_p
H = ctypes.c_void_p  # opaque handles

# name -> (restype, argtypes); every function returns int status unless noted.
SIGNATURES = {
    "mrg8_abi_version": (ctypes.c_int, []),
    "mrg8_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "mrg8_last_error": (ctypes.c_char_p, []),
    "mrg8_init_device": (ctypes.c_int, [ctypes.c_int]),
    "mrg8_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(H)]),
    "mrg8_destroy": (ctypes.c_int, [H]),
    "mrg8_seed_init": (ctypes.c_int, [H, ctypes.c_uint32]),
    "mrg8_get_state": (ctypes.c_int, [H, u32p]),
    "mrg8_set_state": (ctypes.c_int, [H, u32p]),
    "mrg8_jump_ahead": (ctypes.c_int, [H, ctypes.c_uint64]),
    "mrg8_jump_ahead_bits": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_short)]),
    "mrg8_jump_state": (ctypes.c_int, [H, ctypes.c_uint64, u32p]),
    "mrg8_jump_state_bits": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_short), u32p]),
    "mrg8_seed_state": (ctypes.c_int, [ctypes.c_uint32, u32p]),
    "mrg8_advance_state": (ctypes.c_int, [u32p, ctypes.c_uint64, u32p]),
    "mrg8_advance_state_bits": (ctypes.c_int, [u32p, ctypes.POINTER(ctypes.c_short), u32p]),
    "mrg8_generate": (ctypes.c_int, [H, vp, ctypes.c_uint64, ctypes.c_int, vp]),
    "mrg8_generate_shape": (ctypes.c_int, [H, vp, ctypes.c_uint64, ctypes.c_int, vp, ctypes.c_int, ctypes.c_uint64,
                                           ctypes.c_int]),
    "mrg8_tune_fill": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p,
                                      ctypes.POINTER(FillShape)]),
    "mrg8_outer_32": (ctypes.c_int, [H, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mrg8_inner": (ctypes.c_int, [H, vp, ctypes.c_int, ctypes.c_int]),
    "mrg8_state_cpy_DtH": (ctypes.c_int, [H]),
    "mrg8_synchronize": (ctypes.c_int, [H]),
    "mrg8_randint": (ctypes.c_int, [H, vp, ctypes.c_uint64]),
    "mrg8_rand": (ctypes.c_int, [H, vp, ctypes.c_uint64]),
    "mrg8_mrg8dnz2": (ctypes.c_int, [H, vp, ctypes.c_uint64]),
    "mrg8_randint1": (ctypes.c_int, [H, u32p]),
    "mrg8_rand1": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_double)]),
    "mrg8_write_jump_matrix": (ctypes.c_int, [ctypes.c_char_p]),
    "mrg8_jump_matrix": (ctypes.c_int, [u32p, ctypes.c_size_t]),
    "mrg8_read_jump_matrix": (ctypes.c_int, [ctypes.c_char_p, u32p, ctypes.c_size_t]),
    "mrg8_check_jump_matrix": (ctypes.c_int, [ctypes.c_char_p]),
    "mrg8_advance_states": (ctypes.c_int, [ctypes.c_int, vp, vp, ctypes.c_uint32, vp, ctypes.c_uint64, vp]),
    "mrg8_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(H)]),
    "mrg8_batch_destroy": (ctypes.c_int, [H]),
    "mrg8_batch_get_config": (ctypes.c_int, [H, ctypes.POINTER(BatchConfig)]),
    "mrg8_batch_set_config": (ctypes.c_int, [H, ctypes.POINTER(BatchConfig)]),
    "mrg8_batch_set": (ctypes.c_int, [H, ctypes.POINTER(StreamDesc), ctypes.c_uint32]),
    "mrg8_batch_run": (ctypes.c_int, [H, vp]),
    "mrg8_batch_use_graph": (ctypes.c_int, [H, ctypes.c_int]),
    "mrg8_batch_autotune": (ctypes.c_int, [H, vp, ctypes.c_char_p, ctypes.POINTER(BatchConfig)]),
    "mrg8_batch_generate_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(StreamDesc), ctypes.c_uint32,
                                                   ctypes.POINTER(BatchConfig), vp]),
    "mrg8_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ServerConfig), ctypes.c_uint32,
                                          ctypes.POINTER(H)]),
    "mrg8_server_destroy": (ctypes.c_int, [H]),
    "mrg8_server_set_stream": (ctypes.c_int, [H, ctypes.c_uint32, u32p, ctypes.c_uint64]),
    "mrg8_server_set_streams": (ctypes.c_int, [H, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_void_p]),
    "mrg8_server_get_stream": (ctypes.c_int, [H, ctypes.c_uint32, u32p, ctypes.POINTER(ctypes.c_uint64)]),
    "mrg8_server_submit": (ctypes.c_int, [H, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "mrg8_server_progress": (ctypes.c_int, [H, ctypes.POINTER(Completion), ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]),
    "mrg8_server_release": (ctypes.c_int, [H, ctypes.c_uint64]),
    "mrg8_server_stats": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "mrg8_server_counters_get": (ctypes.c_int, [H, ctypes.POINTER(ServerCounters)]),
    "mrg8_server_replay": (ctypes.c_int, [H, ctypes.POINTER(TraceEntry), ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(TraceStats)]),
}

_lib = None


class Mrg8Error(RuntimeError):
    def __init__(self, code: int, func: str, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmrg8_amd.so not built ({LIB_PATH}); run `python -m mrg8_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if rc != OK:
        msg = lib().mrg8_last_error()
        raise Mrg8Error(rc, name, msg.decode() if msg else lib().mrg8_strerror(rc).decode())
    return rc
