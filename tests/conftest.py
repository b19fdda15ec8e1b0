import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The product package on a real device; fails (never skips) when the
    library or the device is missing, so -m gpu runs cannot pass vacuously."""
    import mrg8_amd
    from mrg8_amd import hip

    assert hip.device_count() > 0, "no HIP device visible"
    mrg8_amd.capi.call("mrg8_init_device", 0)
    return mrg8_amd


def wsum(arr: np.ndarray) -> int:
    """Position-weighted checksum sum_i w_i (2i+1) mod 2^64 over the 64-bit
    words of float64 outputs or the zero-extended uint32 outputs."""
    if arr.dtype == np.float64:
        w = arr.view(np.uint64)
    else:
        w = arr.astype(np.uint64)
    idx = np.arange(w.size, dtype=np.uint64)
    return int(np.sum(w * (idx * np.uint64(2) + np.uint64(1)), dtype=np.uint64))
