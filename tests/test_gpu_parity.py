"""GPU parity: the gfx950 kernels against the oracle (plain-C restatement of
original/mrg8.cpp), bit for bit, through the C ABI."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 13579  # the reference's test seed (cuda/tests/test_mrg8_outer.cu:35)
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
NP_DTYPE = {0: np.float64, 1: np.float64, 2: np.uint32}


def say(*a):
    print("[case]", *a, flush=True)


def run_device(gpu, g, n, fmt, pad_elems=0, **shape):
    from mrg8_amd.hip import DeviceBuffer

    dt = NP_DTYPE[fmt]
    esz = np.dtype(dt).itemsize
    buf = DeviceBuffer((n + pad_elems + 1) * esz)
    g.generate(buf.ptr + pad_elems * esz, n, fmt, **shape)
    g.synchronize()
    out = buf.copy_to_host(dt, n, pad_elems * esz)
    buf.free()
    return out


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_c1_full_compare(gpu, oracle_mod, fmt):
    """2^24 outputs from seed 13579 compared in full (BASELINE config 1 shape)."""
    n = 1 << 24
    say("c1", fmt, n)
    g = gpu.mrg8_cuda(SEED)
    got = run_device(gpu, g, n, fmt)
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED), n, fmt)
    bad = np.flatnonzero(got != want) if fmt == 2 else np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, f"first mismatch at {bad[:5]}: got {got[bad[:5]]} want {want[bad[:5]]}"
    assert np.array_equal(g.get_state(), st)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 63, 64, 65, 127, 128, 129, 255, 1000, 4095, 65537, 300001])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_ragged_sizes(gpu, oracle_mod, n, fmt):
    say("ragged", n, fmt)
    g = gpu.mrg8_cuda(SEED + n)
    got = run_device(gpu, g, n, fmt)
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED + n), n, fmt)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    assert np.array_equal(g.get_state(), st)


@pytest.mark.parametrize("pad", [1, 2, 3])
@pytest.mark.parametrize("fmt", [0, 2])
def test_unaligned_outputs(gpu, oracle_mod, pad, fmt):
    """Outputs that are only element-aligned: the waves peel to 1 KiB."""
    n = 100003
    say("unaligned", pad, fmt)
    g = gpu.mrg8_cuda(SEED)
    got = run_device(gpu, g, n, fmt, pad_elems=pad)
    want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), n, fmt)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 129])
@pytest.mark.parametrize("pad", [1, 2, 3])
def test_tiny_unaligned_outputs(gpu, oracle_mod, n, pad):
    """Fills shorter than the alignment peel (only wave 0's peel writes)."""
    for fmt in (0, 1, 2):
        say("tiny-unaligned", n, pad, fmt)
        g = gpu.mrg8_cuda(SEED)
        got = run_device(gpu, g, n, fmt, pad_elems=pad)
        want, st = oracle_mod.generate(oracle_mod.seed_state(SEED), n, fmt)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
        assert np.array_equal(g.get_state(), st)


@pytest.mark.parametrize("fmt", [0, 2])
@pytest.mark.parametrize("off_elems", [1, 2, 63, 64, 65, 127, 128, 129, 255])
@pytest.mark.parametrize("n", [50, 127, 200, 3000, 70001])
def test_peel_to_1kib_boundary(gpu, oracle_mod, fmt, off_elems, n):
    """Destinations at element offsets within a 1 KiB line: wave 0 writes the
    outputs before the first 1 KiB boundary one per lane (up to 63 per round,
    so several rounds), the waves after it start on 1 KiB boundaries; the
    bytes around the destination stay untouched."""
    from mrg8_amd.hip import DeviceBuffer

    esz = 8 if fmt != 2 else 4
    if fmt == 0 and off_elems >= 128:
        pytest.skip("offsets of 128 doubles and more are whole 1 KiB lines")
    say("peel-1kib", fmt, off_elems, n)
    buf = DeviceBuffer(n * esz + 4096)
    buf.fill_byte(0xCD)
    base = (buf.ptr + 1023) // 1024 * 1024
    g = gpu.mrg8_cuda(SEED + off_elems)
    g.generate(base + off_elems * esz, n, fmt, block_threads=256, tile=512)
    g.synchronize()
    host = buf.copy_to_host(np.uint8, n * esz + 4096)
    start = base - buf.ptr + off_elems * esz
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED + off_elems), n, fmt)
    assert host[start:start + n * esz].tobytes() == want.tobytes()
    assert (host[:start] == 0xCD).all() and (host[start + n * esz:] == 0xCD).all()
    assert np.array_equal(g.get_state(), st)
    buf.free()


def test_random_fills_default_shape(gpu, oracle_mod):
    """60 seeded random cases with the library's own launch shape: sizes from
    1 to 2^22 (log-uniform, so every block-size and tile rule of the default
    shape is taken), every format, destinations at random element offsets
    within a 4 KiB line, random states (a few with zero words); outputs, the
    bytes around them and the final state against the oracle."""
    from mrg8_amd.hip import DeviceBuffer

    rng = np.random.default_rng(2024)
    p = 2**31 - 1
    buf = DeviceBuffer((1 << 22) * 8 + 16384)
    try:
        base = (buf.ptr + 4095) // 4096 * 4096
        for case in range(60):
            n = int(np.exp(rng.uniform(0, np.log(1 << 22))))
            fmt = int(rng.integers(0, 3))
            esz = 4 if fmt == 2 else 8
            off = int(rng.integers(0, 4096 // esz)) * esz
            st = rng.integers(1, p + 1, size=8, dtype=np.int64).astype(np.uint32)
            if case % 10 == 3:
                st[rng.integers(0, 8, size=3)] = 0
            say("random fill", case, n, fmt, off, st.tolist())
            buf.fill_byte(0xCD)
            g = gpu.mrg8_cuda(1)
            g.set_state(st)
            g.generate(base + off, n, fmt)
            g.synchronize()
            lo = base - buf.ptr + off
            host = buf.copy_to_host(np.uint8, lo + n * esz + 64)
            want, st_after = oracle_mod.generate(st, n, fmt)
            assert host[lo:lo + n * esz].tobytes() == want.tobytes(), case
            assert (host[:lo] == 0xCD).all() and (host[lo + n * esz:] == 0xCD).all(), case
            assert np.array_equal(g.get_state(), st_after), case
    finally:
        buf.free()


@pytest.mark.parametrize("fmt,per_lane", [(0, 1), (0, 2), (0, 4), (1, 4), (2, 1), (2, 2), (2, 4)])
def test_outputs_per_lane(gpu, oracle_mod, fmt, per_lane):
    n = 777_777
    say("per-lane", fmt, per_lane)
    g = gpu.mrg8_cuda(SEED)
    got = run_device(gpu, g, n, fmt, pad_elems=1, block_threads=512, tile=8192, per_lane=per_lane)
    want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), n, fmt)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("n", [1, 100, 128, 129, 255, 256, 257, 511, 1000, 4095, 65537])
@pytest.mark.parametrize("fmt,pad", [(0, 0), (0, 1), (1, 1)])
def test_two_store_layout_ragged(gpu, oracle_mod, n, fmt, pad):
    """Doubles with 4 outputs per lane (two 1 KiB stores per iteration):
    partial last iterations and peeled starts."""
    say("split-ragged", n, fmt, pad)
    g = gpu.mrg8_cuda(SEED + n)
    got = run_device(gpu, g, n, fmt, pad, block_threads=256, tile=512, per_lane=4)
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED + n), n, fmt)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    assert np.array_equal(g.get_state(), st)


# tiles of n and above (2^64 - 1 included) give one wave the whole fill
@pytest.mark.parametrize("block,tile", [(64, 128), (256, 1024), (512, 4096), (1024, 65536), (256, 1 << 20),
                                        (512, 1_000_003), (256, 2**64 - 1)])
def test_launch_shapes_identical(gpu, oracle_mod, block, tile):
    n = 1_000_003
    say("shape", block, tile)
    g = gpu.mrg8_cuda(SEED)
    got = run_device(gpu, g, n, 0, block_threads=block, tile=tile)
    want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), n, 0)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("block", [64, 256, 512, 1024])
@pytest.mark.parametrize("fmt,pad", [(0, 1), (1, 0)])
def test_xcd_runs_cover_every_wave(gpu, oracle_mod, block, fmt, pad):
    """Double fills of 64 or more waves per CU give each XCD runs of
    consecutive blocks' output.  With 256-output waves the grid has whole
    spans of runs plus a remainder of blocks in dispatch order; the buffer is
    poisoned first, so a part written twice or never shows up."""
    import torch
    from mrg8_amd.hip import DeviceBuffer

    cu = torch.cuda.get_device_properties(0).multi_processor_count
    n = cu * 64 * 256 * 19 // 16 + 3
    say("xcd-runs", block, fmt, pad, n)
    buf = DeviceBuffer((n + pad) * 8)
    try:
        buf.fill_byte(0xA5)
        g = gpu.mrg8_cuda(SEED)
        g.jump_ahead(4242)
        g.generate(buf.ptr + pad * 8, n, fmt, block_threads=block, tile=256)
        g.synchronize()
        got = buf.copy_to_host(np.float64, n, pad * 8)
        s0 = oracle_mod.jump(oracle_mod.seed_state(SEED), 4242)
        want, st = oracle_mod.generate(s0, n, fmt)
        bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
        assert bad.size == 0, f"{bad.size} outputs differ, first at {bad[:4]}"
        assert np.array_equal(g.get_state(), st)
    finally:
        buf.free()


def test_beyond_2e32_outputs(gpu, oracle_mod):
    """One call of 2^32 + 4096 words (16 GiB): 64-bit counts and offsets
    throughout (the reference takes an int n and jumps 32 bits on the
    device).  Windows around 2^31 and 2^32 and the final state are checked
    against the oracle's jump."""
    from mrg8_amd.hip import DeviceBuffer

    n = (1 << 32) + 4096
    say("beyond-2^32", n)
    buf = DeviceBuffer(n * 4)
    try:
        g = gpu.mrg8_cuda(SEED)
        g.generate(buf.ptr, n, 2)
        g.synchronize()
        s0 = oracle_mod.seed_state(SEED)
        for start, count in ((0, 4096), ((1 << 31) - 2048, 4096), ((1 << 32) - 2048, 6144)):
            got = buf.copy_to_host(np.uint32, count, start * 4)
            want, _ = oracle_mod.generate(oracle_mod.jump(s0, start), count, 2)
            assert np.array_equal(got, want), start
        assert np.array_equal(g.get_state(), oracle_mod.jump(s0, n))
    finally:
        buf.free()


def test_consecutive_calls_continue(gpu, oracle_mod):
    say("consecutive")
    g = gpu.mrg8_cuda(SEED)
    parts = [run_device(gpu, g, n, 2) for n in (5, 1000, 77, 4096, 123457)]
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED), sum(p.size for p in parts), 2)
    assert np.array_equal(np.concatenate(parts), want)
    assert np.array_equal(g.get_state(), st)


def test_set_state_and_jump(gpu, oracle_mod):
    rng = np.random.default_rng(7)
    for trial in range(20):
        st = rng.integers(1, 2**31 - 1, size=8, dtype=np.uint64).astype(np.uint32)
        step = int(rng.integers(0, 2**63, dtype=np.uint64))
        say("jump", trial, step)
        g = gpu.mrg8_cuda(1)
        g.set_state(st)
        g.jump_ahead(step)
        got = run_device(gpu, g, 5000, 2)
        want, _ = oracle_mod.generate(oracle_mod.jump(st, step), 5000, 2)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("words", [
    [0, 0, 0, 0, 0, 0, 0, 5],
    [0, 2147483647, 0, 0, 0, 0, 0, 0],
    [2147483647] * 8,
    [0] * 8,
    [1, 0, 1, 0, 1, 0, 1, 0],
])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_noncanonical_states(gpu, oracle_mod, words, fmt):
    """Zero and p words in the state (both allowed by set_state) keep the
    reference's representation through the kernels' jump and generation."""
    say("noncanonical", words, fmt)
    w = np.array(words, dtype=np.uint32)
    for offset in (0, 1, 3, 9, 100):
        g = gpu.mrg8_cuda(1)
        g.set_state(w)
        g.jump_ahead(offset)
        n = 3000
        # tiny tiles: every wave after the first starts with a device jump
        got = run_device(gpu, g, n, fmt, block_threads=64, tile=64)
        want, st = oracle_mod.generate(oracle_mod.jump(w, offset), n, fmt)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), offset
        assert np.array_equal(g.get_state(), st)


def test_device_jump_small_offsets(gpu, oracle_mod):
    """Waves that start at small offsets (1..8 positions) exercise the shift rows of A, A^2, A^4."""
    for n in range(129, 129 + 16):
        say("small-offset", n)
        g = gpu.mrg8_cuda(SEED)
        got = run_device(gpu, g, n, 2, block_threads=64, tile=128)
        want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), n, 2)
        assert np.array_equal(got, want)


def test_drop_in_outer_inner(gpu, oracle_mod):
    from mrg8_amd.hip import DeviceBuffer

    n = 1 << 20
    say("drop-in")
    buf = DeviceBuffer(n * 8)
    calls = {
        "outer": lambda g: g.mrg8_outer_32(buf, n, 64, 65536),
        "outer, TNUM only (BS 64)": lambda g: g.mrg8_outer_32(buf, n, TNUM=65536),
        "outer, library shape": lambda g: g.mrg8_outer_32(buf, n),
        "inner": lambda g: g.mrg8_inner(buf, n, 65536),
        "inner, library shape": lambda g: g.mrg8_inner(buf, n, 0),
    }
    for call, fn in calls.items():
        g = gpu.mrg8_cuda(SEED)
        fn(g)
        got = buf.copy_to_host(np.float64, n)
        want, st = oracle_mod.generate(oracle_mod.seed_state(SEED), n, 1)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), call
        g.state_cpy_DtH()
        assert np.array_equal(g.get_state(), st)
    buf.free()


def test_host_forms(gpu, oracle_mod):
    say("host-forms")
    g = gpu.mrg8_cuda(SEED)
    a = g.randint(1000)
    b = g.rand(2000)
    c = g.mrg8dnz2(3000)
    d = g.randint()
    e = g.rand()
    want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), 6002, 2)
    rn = 1.0 / 2147483647.0
    assert np.array_equal(a, want[:1000])
    assert np.array_equal(b, want[1000:3000].astype(np.float64) * rn)
    assert np.array_equal(c, want[3000:6000].astype(np.float64) * rn)
    assert d == int(want[6000])
    assert e == float(want[6001]) * rn


def test_host_forms_many_chunks(gpu, oracle_mod):
    """Host-memory forms larger than the staging chunks (32 MiB): 5 chunks of
    doubles and 2.5 of words, plus a device call in between; the sequence
    continues across all of them."""
    say("host-forms-chunks")
    g = gpu.mrg8_cuda(SEED + 1)
    nd = 5 * (4 << 20) + 12345   # 4 Mi doubles per chunk
    nw = (8 << 20) * 2 + 777     # 8 Mi words per chunk
    a = g.rand(nd)
    b = g.randint(nw)
    c = run_device(gpu, g, 1001, 0)
    d = g.rand(3)
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED + 1), nd, 0)
    assert np.array_equal(a.view(np.uint64), want.view(np.uint64))
    want, st = oracle_mod.generate(st, nw, 2)
    assert np.array_equal(b, want)
    want, st = oracle_mod.generate(st, 1001, 0)
    assert np.array_equal(c.view(np.uint64), want.view(np.uint64))
    want, st = oracle_mod.generate(st, 3, 0)
    assert np.array_equal(d.view(np.uint64), want.view(np.uint64))
    assert np.array_equal(g.get_state(), st)


def test_c2_sampled_and_hash(gpu, oracle_mod):
    """2^28 outputs in u32 and f64: 1e4 sampled indices against jump+1 output,
    and the FNV-1a hash of the whole sequence against the oracle's and the
    compiled reference's."""
    from mrg8_amd.hip import DeviceBuffer

    n = 1 << 28
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.integers(0, n, size=10_000, dtype=np.uint64))
    s0 = oracle_mod.seed_state(SEED)
    for fmt in (2, 0):
        say("c2", fmt)
        dt = NP_DTYPE[fmt]
        buf = DeviceBuffer(n * np.dtype(dt).itemsize)
        g = gpu.mrg8_cuda(SEED)
        g.generate(buf, n, fmt)
        g.synchronize()
        got = buf.copy_to_host(dt, n)
        buf.free()
        for i in idx:
            v, _ = oracle_mod.generate(oracle_mod.jump(s0, int(i)), 1, fmt)
            assert got[int(i)].tobytes() == v[0].tobytes(), int(i)
        h, st = oracle_mod.hash_stream(s0, n, fmt)
        assert oracle_mod.fnv1a(got) == h
        # the compiled reference's own hash and final state for this sequence
        key = "randint_2^28" if fmt == 2 else "rand_2^28"
        assert f"{oracle_mod.fnv1a(got):016x}" == GOLD["hash"][key]["fnv1a64"]
        assert g.get_state().tolist() == GOLD["hash"][key]["state"]
        assert np.array_equal(g.get_state(), st)
        del got


def test_fill_autotune_keeps_outputs(gpu, oracle_mod, tmp_path):
    """The tuned shape becomes the default for that size class; outputs stay bit-exact."""
    cache = str(tmp_path / "fill_tune.txt")
    for fmt in (0, 2):
        say("tune-fill", fmt)
        shape = gpu.tune_fill(fmt, 1 << 22, cache_path=cache)
        assert shape["block_threads"] in (256, 512, 1024) and shape["outputs_per_lane"] in (1, 2, 4, 16, 32)
        again = gpu.tune_fill(fmt, 1 << 22, cache_path=cache)  # served from the cache file
        assert again == shape
        g = gpu.mrg8_cuda(SEED)
        n = (1 << 22) + 12345
        got = run_device(gpu, g, n, fmt)
        want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), n, fmt)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    assert len(open(cache).read().splitlines()) == 2


def test_threads_with_own_handles_and_streams(gpu, oracle_mod):
    """Eight host threads, each with its own generator and HIP stream, fill
    concurrently; every result matches the oracle."""
    import threading

    from mrg8_amd import hip

    n = 300_007
    results, errors = {}, []

    def work(t):
        try:
            s = hip.Stream()
            buf = hip.DeviceBuffer(n * 8)
            g = gpu.mrg8_cuda(1000 + t)
            g.jump_ahead(t * 12345)
            for _ in range(3):  # the third call's outputs are kept
                g.generate(buf, n, t % 3, stream=s)
            s.synchronize()
            results[t] = buf.copy_to_host(NP_DTYPE[t % 3], n)
            buf.free()
            s.destroy()
        except Exception as e:  # pragma: no cover - reported below
            errors.append(repr(e))

    say("threads")
    ths = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors
    for t in range(8):
        start = oracle_mod.jump(oracle_mod.seed_state(1000 + t), t * 12345 + 2 * n)
        want, _ = oracle_mod.generate(start, n, t % 3)
        assert np.array_equal(results[t].view(np.uint8), want.view(np.uint8)), t


def test_one_handle_on_two_streams(gpu, oracle_mod):
    """Successive calls on two streams continue the sequence; synchronize
    waits for both, and a handle destroyed with work in flight leaves that
    work to finish on its stream."""
    from mrg8_amd import hip

    n = 1 << 20
    say("two streams")
    a, b = hip.Stream(), hip.Stream()
    bufs = [hip.DeviceBuffer(n * 8) for _ in range(3)]
    g = gpu.mrg8_cuda(SEED)
    g.generate(bufs[0], n, 0, stream=a)
    g.generate(bufs[1], n, 0, stream=b)
    g.synchronize()
    g.generate(bufs[2], n, 0, stream=b)
    g.close()  # no wait: the launch completes on stream b
    b.synchronize()
    want, _ = oracle_mod.generate(oracle_mod.seed_state(SEED), 3 * n, 0)
    for i, buf in enumerate(bufs):
        got = buf.copy_to_host(np.float64, n)
        assert np.array_equal(got.view(np.uint64), want[i * n : (i + 1) * n].view(np.uint64)), i
        buf.free()
    a.destroy()
    b.destroy()


def test_generate_inside_a_hip_graph(gpu, oracle_mod):
    """A generate call captured into a HIP graph (torch.cuda.graph) replays
    the positions it had at capture: each replay rewrites the same outputs,
    identical to the eager call; the handle's state advanced once, at
    capture."""
    import torch

    n = 1 << 20
    g = gpu.mrg8_cuda(SEED)
    out = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    side = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    say("graph-capture")
    with torch.cuda.graph(graph, stream=side):
        g.generate(out, n, 0, stream=side)
    want, st = oracle_mod.generate(oracle_mod.seed_state(SEED), n, 0)
    assert np.array_equal(g.get_state(), st)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint64), want.view(np.uint64)), rep


@pytest.mark.parametrize("fmt", [0, 1])
def test_c2_statistics_like_check_rand(gpu, fmt):
    """The reference's only check of its generators is statistical
    (common/rng_test.h:29-47: mean and standard deviation of the outputs).
    2^28 doubles: mean 1/2 and standard deviation 1/sqrt(12) within about
    4 standard errors, and every value inside its interval ((0, 1] for
    x / p, [0, 1) for (x - 1) / p)."""
    import torch

    n = 1 << 28
    say("check-rand", fmt)
    t = torch.empty(n, dtype=torch.float64, device="cuda:0")
    g = gpu.mrg8_cuda(SEED)
    g.generate(t, n, fmt, stream=torch.cuda.current_stream())
    mean = t.mean().item()
    std = t.std().item()
    lo, hi = t.min().item(), t.max().item()
    del t
    se = (1 / 12) ** 0.5 / n ** 0.5
    assert abs(mean - 0.5) < 4 * se, mean
    assert abs(std - (1 / 12) ** 0.5) < 4 * se, std  # the sample sd of U(0,1) varies by about 0.45 se
    if fmt == 0:
        assert lo > 0.0 and hi <= 1.0
    else:
        assert lo >= 0.0 and hi < 1.0


def test_benchmark_driver_harness(gpu, capsys):
    """The counterpart of cuda/tests/test_mrg8_outer.cu / _inner.cu prints the
    reference's EVALUATION lines for a short sweep."""
    from mrg8_amd import harness

    say("harness")
    assert harness.main(["1", "--tnum-max", "2048", "--auto"]) == 0
    assert harness.main(["1", "--tnum-max", "1024", "--inner"]) == 0
    out = capsys.readouterr().out
    lines = [l for l in out.splitlines() if l.startswith("EVALUATION")]
    # outer: TNUM 1024 x BS 32..1024 (6) + TNUM 2048 (6) + auto; inner: TNUM 1024
    assert len(lines) == 6 + 6 + 1 + 1, lines
    f = lines[0].split(", ")
    assert f[1] == "MRG8_HIP_outer" and f[2] == "1024" and f[3] == "32" and f[4] == str(1 << 20)
    assert float(f[5]) > 0 and float(f[6]) > 0
    assert lines[-1].split(", ")[3] == ""  # inner: empty BS field
    assert "Arithmetic mean: 0.49" in out or "Arithmetic mean: 0.50" in out
