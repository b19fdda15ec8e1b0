"""Synthetic inputs for the MRG8 benchmark configurations (fixed seeds, numpy PCG64).

    single  : one stream from seed 13579 (the reference's test seed,
              cuda/tests/test_mrg8_outer.cu:35), n outputs
    batched : 4096 streams, states uniform in [1, 2^31-2]^8, lengths
              floor(2^(10 + 12u)) (log-uniform in [2^10, 2^22)), float64
    jumps   : 65536 streams from one seed, each jumped by a uniform step in
              [0, 2^63), then 256 uint32 outputs
    served  : 16384 resident streams, Poisson arrivals, sizes floor(2^(8 + 18u))
              float64, stream ids uniform
"""
from __future__ import annotations

import numpy as np

SEED = 13579


def batched(n_streams: int = 4096, lo_exp: int = 10, hi_exp: int = 22, seed: int = 42):
    rng = np.random.default_rng(seed)
    states = rng.integers(1, 2**31 - 1, size=(n_streams, 8), dtype=np.int64).astype(np.uint32)
    u = rng.random(n_streams)
    lengths = np.floor(2.0 ** (lo_exp + (hi_exp - lo_exp) * u)).astype(np.uint64)
    return states, lengths


def jump_steps(n_streams: int = 65536, seed: int = 43):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**63, size=n_streams, dtype=np.uint64)


def served(n_requests: int = 100_000, n_streams: int = 16384, rate_per_s: float = 50_000.0,
           lo_exp: int = 8, hi_exp: int = 26, seed: int = 44):
    """Returns (stream_states[n_streams, 8], arrivals_s, stream_ids, sizes)."""
    rng = np.random.default_rng(seed)
    states = rng.integers(1, 2**31 - 1, size=(n_streams, 8), dtype=np.int64).astype(np.uint32)
    gaps = rng.exponential(1.0 / rate_per_s, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    ids = rng.integers(0, n_streams, size=n_requests, dtype=np.int64).astype(np.uint32)
    u = rng.random(n_requests)
    sizes = np.floor(2.0 ** (lo_exp + (hi_exp - lo_exp) * u)).astype(np.uint64)
    return states, arrivals, ids, sizes


def mean_served_size(lo_exp: int = 8, hi_exp: int = 26) -> float:
    """E[floor(2^(lo + (hi-lo)u))] ~ (2^hi - 2^lo) / ((hi-lo) ln 2)."""
    return (2.0**hi_exp - 2.0**lo_exp) / ((hi_exp - lo_exp) * np.log(2.0))
