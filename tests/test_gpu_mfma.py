"""GPU parity of the matrix-core fill (k_fill_mfma: int8 MFMAs over byte
digits, outputs_per_lane 16 or 32) against the oracle (plain-C restatement
of original/mrg8.cpp), bit for bit, through the C ABI: tile shapes, block
sizes, runs per wave, unaligned destinations with their short head and tail
runs, states with zero and p words, the all-zero state, long fills checked
in windows, and the shape errors."""
import numpy as np
import pytest

from test_gpu_parity import SEED, run_device, say

pytestmark = pytest.mark.gpu

TILE = {16: 1024, 32: 2048}  # outputs per tile: 16 columns of 4 x per_lane


def check(oracle_mod, got, state, n, fmt, g=None):
    want, st = oracle_mod.generate(state, n, fmt)
    bad = np.flatnonzero(got.view(np.uint8) != want.view(np.uint8))
    assert bad.size == 0, f"first differing byte {bad[:4]}"
    if g is not None:
        assert np.array_equal(g.get_state(), st)


@pytest.mark.parametrize("per_lane", [16, 32])
@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("n,pad", [(2048 * 3 + 5, 0), (65536 + 777, 1), ((1 << 20) + 123, 3), (4096, 0), (5000, 2)])
def test_matches_oracle(gpu, oracle_mod, per_lane, fmt, n, pad):
    say("mfma", per_lane, fmt, n, pad)
    g = gpu.mrg8_cuda(SEED + n)
    got = run_device(gpu, g, n, fmt, pad_elems=pad, per_lane=per_lane)
    check(oracle_mod, got, oracle_mod.seed_state(SEED + n), n, fmt, g)


@pytest.mark.parametrize("per_lane", [16, 32])
@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("tiles_per_wave", [1, 3, 4, 33])
def test_block_sizes_and_runs_per_wave(gpu, oracle_mod, per_lane, block, tiles_per_wave):
    """Every wave makes tiles_per_wave consecutive tiles (the last wave
    fewer), moving its 16 column states with the A^(tile) product."""
    n = 300_007
    fmt = 2 if block != 1024 else 0
    say("mfma-shape", per_lane, block, tiles_per_wave)
    g = gpu.mrg8_cuda(SEED)
    got = run_device(gpu, g, n, fmt, pad_elems=1, per_lane=per_lane, block_threads=block,
                     tile=TILE[per_lane] * tiles_per_wave)
    check(oracle_mod, got, oracle_mod.seed_state(SEED), n, fmt, g)


@pytest.mark.parametrize("words", [
    [0, 0, 0, 0, 0, 0, 0, 5],
    [0, 2147483647, 0, 0, 0, 0, 0, 0],
    [2147483647] * 8,
    [1, 0, 1, 0, 1, 0, 1, 0],
    [2147483646, 1, 2147483647, 0, 3, 2147483647, 2, 1],
])
@pytest.mark.parametrize("per_lane", [16, 32])
def test_noncanonical_states(gpu, oracle_mod, words, per_lane):
    """Zero and p words reach the tile states (column 0 of a wave's first tile
    is the tile state itself) and the digit operands."""
    w = np.array(words, dtype=np.uint32)
    for offset in (0, 1, 5, 2048):
        for fmt in (0, 2):
            say("mfma-noncanonical", words, per_lane, offset, fmt)
            g = gpu.mrg8_cuda(1)
            g.set_state(w)
            g.jump_ahead(offset)
            n = 40_000
            got = run_device(gpu, g, n, fmt, per_lane=per_lane, tile=TILE[per_lane] * 2)
            check(oracle_mod, got, oracle_mod.jump(w, offset), n, fmt, g)


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_all_zero_state(gpu, oracle_mod, fmt):
    """The all-zero state stays zero: the fill leaves it to the vector kernel."""
    say("mfma-zero", fmt)
    g = gpu.mrg8_cuda(1)
    g.set_state(np.zeros(8, dtype=np.uint32))
    n = 100_000
    got = run_device(gpu, g, n, fmt, per_lane=32)
    check(oracle_mod, got, np.zeros(8, dtype=np.uint32), n, fmt, g)


@pytest.mark.parametrize("per_lane,fmt", [(16, 0), (32, 2), (0, 2)])
def test_long_fill_windows(gpu, oracle_mod, per_lane, fmt):
    """2^26 outputs (per_lane 0: the default, the matrix-core fill for words)
    from a jumped state, checked in windows around tile and wave boundaries
    and at the end, with the final state."""
    from mrg8_amd.hip import DeviceBuffer

    n = 1 << 26
    step = 123_456_789_012
    say("mfma-long", per_lane, fmt)
    esz = 4 if fmt == 2 else 8
    dt = np.uint32 if fmt == 2 else np.float64
    s0 = oracle_mod.jump(oracle_mod.seed_state(SEED), step)
    buf = DeviceBuffer(n * esz + 64)
    try:
        g = gpu.mrg8_cuda(SEED)
        g.jump_ahead(step)
        g.generate(buf.ptr + 8, n, fmt, per_lane=per_lane)
        g.synchronize()
        for start in (0, 1000, 8190, (1 << 20) - 7, (1 << 25) + 2040, n - 5000):
            count = min(5000, n - start)
            got = buf.copy_to_host(dt, count, 8 + start * esz)
            want, _ = oracle_mod.generate(oracle_mod.jump(s0, start), count, fmt)
            assert got.tobytes() == want.tobytes(), start
        assert np.array_equal(g.get_state(), oracle_mod.jump(s0, n))
    finally:
        buf.free()


def test_shape_errors(gpu):
    from mrg8_amd.capi import Mrg8Error
    from mrg8_amd.hip import DeviceBuffer

    buf = DeviceBuffer(1 << 20)
    try:
        g = gpu.mrg8_cuda(SEED)
        for kw in (dict(per_lane=16, block_threads=128), dict(per_lane=32, tile=1024), dict(per_lane=16, tile=1000),
                   dict(per_lane=8)):
            with pytest.raises(Mrg8Error):
                g.generate(buf.ptr, 100_000, 2, **kw)
    finally:
        buf.free()


@pytest.mark.parametrize("fmt,per_lane", [(2, 0), (0, 16)])
def test_inside_a_hip_graph(gpu, oracle_mod, fmt, per_lane):
    """The matrix-core fill captured into a HIP graph (its tables live in the
    device context, so the capture records one kernel node and no
    allocation): every replay rewrites the outputs of the positions the
    handle had at capture, identical to the oracle."""
    import torch

    n = (1 << 22) + 333
    dt = torch.int32 if fmt == 2 else torch.float64
    g = gpu.mrg8_cuda(SEED)
    g.jump_ahead(77)
    out = torch.zeros(n, dtype=dt, device="cuda:0")
    side = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    say("mfma-graph", fmt, per_lane)
    with torch.cuda.graph(graph, stream=side):
        g.generate(out, n, fmt, stream=side, per_lane=per_lane)
    s0 = oracle_mod.jump(oracle_mod.seed_state(SEED), 77)
    want, st = oracle_mod.generate(s0, n, fmt)
    assert np.array_equal(g.get_state(), st)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert out.cpu().numpy().tobytes() == want.tobytes(), rep
