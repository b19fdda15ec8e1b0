"""mrg8_batch_generate_noalloc (the entry point that allocates nothing)
against the existing batched entry point (mrg8_batch_set + mrg8_batch_run)
through the C ABI: identical bytes on two calls in a row, the second with a
smaller input, for the default one-tile-per-wave shape and a persistent grid;
segments of mixed formats, offsets and lengths packed at odd element offsets."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ESZ = {0: 8, 1: 8, 2: 4}


def say(*a):
    print("[case]", *a, flush=True)


def make_segments(rng, n, max_log2):
    states = rng.integers(1, 2**31 - 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    offsets = [0 if i % 3 == 0 else int(rng.integers(0, 2**62)) for i in range(n)]
    counts = [int(2 ** rng.uniform(0, max_log2)) for _ in range(n)]
    fmts = [int(rng.integers(0, 3)) for _ in range(n)]
    return states, offsets, counts, fmts


def place(base, counts, fmts):
    """Byte addresses of the segments, packed back to back (element-aligned)."""
    addr, out = base + 4, []
    for c, f in zip(counts, fmts):
        e = ESZ[f]
        addr = (addr + e - 1) // e * e
        out.append(addr)
        addr += c * e
    return out, addr - base


def descs(capi, states, offsets, counts, fmts, addrs):
    arr = (capi.StreamDesc * max(1, len(counts)))()
    for i in range(len(counts)):
        d = arr[i]
        d.state = (ctypes.c_uint32 * 8)(*[int(x) for x in states[i]])
        d.offset, d.count, d.out, d.format, d.reserved = offsets[i], counts[i], addrs[i], fmts[i], 0
    return arr


@pytest.mark.parametrize("shape", [None, (256, 4, 4096), (512, 0, 256), (512, -2, 2048)],
                         ids=["default", "persistent", "tiles-256", "address-order"])
def test_noalloc_equals_batch_run(gpu, oracle_mod, shape):
    from mrg8_amd import capi, hip

    rng = np.random.default_rng(2024)
    b = gpu.Batch(max_streams=512)
    if shape:
        b.set_config(*shape)
    cfg = capi.BatchConfig(*shape) if shape else None
    for call, (n, max_log2) in enumerate(((400, 17), (37, 12))):  # the second call is smaller
        states, offsets, counts, fmts = make_segments(rng, n, max_log2)
        _, nbytes = place(0, counts, fmts)
        ref, got = hip.DeviceBuffer(nbytes + 64), hip.DeviceBuffer(nbytes + 64)
        ref.fill_byte(0xA5)
        got.fill_byte(0xA5)
        ra, _ = place(ref.ptr, counts, fmts)
        ga, _ = place(got.ptr, counts, fmts)
        say("noalloc", shape, "call", call, n, "segments", sum(counts), "outputs")
        b.set([(states[i], offsets[i], counts[i], ra[i], fmts[i]) for i in range(n)])
        b.run()
        arr = descs(capi, states, offsets, counts, fmts, ga)
        capi.call("mrg8_batch_generate_noalloc", 0, arr, ctypes.c_uint32(n),
                  ctypes.byref(cfg) if cfg is not None else None, None)
        hip.synchronize()
        hr = ref.copy_to_host(np.uint8, nbytes + 64)
        hg = got.copy_to_host(np.uint8, nbytes + 64)
        assert np.array_equal(hr, hg), np.flatnonzero(hr != hg)[:8]
        # and the oracle, for a few segments of each call
        for i in (0, n // 2, n - 1):
            want, _ = oracle_mod.generate(oracle_mod.jump(states[i], offsets[i]), counts[i], fmts[i])
            lo = ga[i] - got.ptr
            assert hg[lo:lo + counts[i] * ESZ[fmts[i]]].tobytes() == want.tobytes(), (call, i)
        ref.free()
        got.free()
    b.close()


def test_noalloc_limits(gpu):
    from mrg8_amd import capi, hip

    buf = hip.DeviceBuffer(1 << 20)
    st = np.arange(1, 9, dtype=np.uint32).reshape(1, 8)
    cfg = capi.BatchConfig(256, 0, 256)
    # 2^30 outputs in 256-output tiles: a 16 MiB tile map, beyond the 4 MiB area
    arr = descs(capi, st, [0], [1 << 30], [0], [buf.ptr])
    rc = capi.lib().mrg8_batch_generate_noalloc(0, arr, 1, ctypes.byref(cfg), None)
    assert rc == capi.ERR_CAPACITY
    assert b"scratch area" in capi.lib().mrg8_last_error()
    arr = descs(capi, st, [0], [16], [0], [buf.ptr])
    assert capi.lib().mrg8_batch_generate_noalloc(99, arr, 1, None, None) == capi.ERR_ARG
    bad = capi.BatchConfig(100, 0, 256)
    assert capi.lib().mrg8_batch_generate_noalloc(0, arr, 1, ctypes.byref(bad), None) == capi.ERR_ARG
    assert capi.lib().mrg8_batch_generate_noalloc(0, arr, 0, None, None) == capi.OK  # nothing to do
    buf.free()
