"""The benchmark's write probe (libmrg8_probe.so) writes d[i] = i over the
whole buffer in each of its store patterns, including the XCD runs and the
blocks past the last whole span of runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,tile", [(0, 0), (1, 128), (1, 2048), (2, 128), (2, 2048)])
def test_write_probe_covers_buffer(gpu, mode, tile):
    from mrg8_amd import hip, probe

    n = 3_000_017
    print("[case] probe", mode, tile, flush=True)
    buf = hip.DeviceBuffer(n * 8)
    try:
        buf.fill_byte(0xA5)
        probe.write(buf.ptr, n, mode, tile)
        hip.synchronize()
        got = buf.copy_to_host(np.float64, n)
        assert np.array_equal(got, np.arange(n, dtype=np.float64))
    finally:
        buf.free()
