"""The rocRAND comparators (the counterpart of the reference's cuRAND drivers,
cuda/tests/test_curand_*.cu) run on the device: every generator reports
positive rates for both timing methods and leaves uniform doubles in the
buffer.  Different generators make different sequences, so only the range and
the first two moments are checked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["xorwow", "mrg32k3a", "mtgp32", "philox4_32_10", "mt19937"])
def test_rocrand_comparator_runs(gpu, name):
    from mrg8_amd import comparators, hip

    n = 1 << 20
    print("[case] rocrand", name, flush=True)
    buf = hip.DeviceBuffer(n * 8)
    try:
        r = comparators.time_generator(name, buf, n, iters=3)
        assert r["reseed_ms"] > 0 and r["steady_ms"] > 0
        assert r["reseed_gsps"] > 0 and r["steady_gsps"] > 0
        x = buf.copy_to_host(np.float64, n)
    finally:
        buf.free()
    assert np.isfinite(x).all() and (x >= 0.0).all() and (x <= 1.0).all()
    assert abs(x.mean() - 0.5) < 0.01 and abs(x.var() - 1.0 / 12.0) < 0.005


def test_rocrand_rates_table(gpu):
    from mrg8_amd import comparators

    rates = comparators.rocrand_rates(1 << 18, iters=2, names=["philox4_32_10", "xorwow"])
    assert set(rates) == {"philox4_32_10", "xorwow"}
    for v in rates.values():
        assert "error" not in v and v["steady_gsps"] > 0
