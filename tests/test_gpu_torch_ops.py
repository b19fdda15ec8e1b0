"""GPU parity of the PyTorch caller (torch HIP tensors filled through the C ABI)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_generator_continues_and_matches_oracle(gpu, oracle_mod):
    import torch

    from mrg8_amd import torch_ops

    g = torch_ops.Generator(13579, device="cuda:0")
    s0 = gpu.seed_state(13579)
    say("torch-generator")
    a = g.rand(100003)
    b = g.rand(4096, fmt=torch_ops.F64_01)
    c = g.randint(777)
    d = g.randint(64, dtype=torch.uint32)
    torch.cuda.synchronize()
    want_a, st = oracle_mod.generate(s0, 100003, 0)
    want_b, st = oracle_mod.generate(st, 4096, 1)
    want_c, st = oracle_mod.generate(st, 777, 2)
    want_d, st = oracle_mod.generate(st, 64, 2)
    assert np.array_equal(a.cpu().numpy().view(np.uint64), want_a.view(np.uint64))
    assert np.array_equal(b.cpu().numpy().view(np.uint64), want_b.view(np.uint64))
    assert np.array_equal(c.cpu().numpy().astype(np.uint32), want_c)
    assert np.array_equal(d.view(torch.int32).cpu().numpy().view(np.uint32), want_d)
    assert np.array_equal(g.get_state(), st)


def test_fill_op_on_side_stream(gpu, oracle_mod):
    import torch

    from mrg8_amd import torch_ops

    rng = np.random.default_rng(3)
    state = [int(x) for x in rng.integers(1, 2**31 - 1, size=8)]
    side = torch.cuda.Stream()
    out = torch.full((3, 50001), -1.0, dtype=torch.float64, device="cuda:0")
    words = torch.zeros(12345, dtype=torch.int32, device="cuda:0")
    say("torch-fill-op")
    with torch.cuda.stream(side):
        torch.ops.mrg8.fill_(out, state, 2**62 + 5, torch_ops.F64)
        torch.ops.mrg8.fill_(words, state, 0, torch_ops.U32)
    side.synchronize()
    start = oracle_mod.jump(np.array(state, dtype=np.uint32), 2**62 + 5)
    want, _ = oracle_mod.generate(start, out.numel(), 0)
    assert np.array_equal(out.cpu().numpy().reshape(-1).view(np.uint64), want.view(np.uint64))
    want_w, _ = oracle_mod.generate(np.array(state, dtype=np.uint32), words.numel(), 2)
    assert np.array_equal(words.cpu().numpy().view(np.uint32), want_w)
    with pytest.raises(TypeError):
        torch.ops.mrg8.fill_(torch.empty(4, dtype=torch.float32, device="cuda:0"), state, 0, torch_ops.F64)
    with pytest.raises(ValueError):
        torch.ops.mrg8.fill_(out[:, ::2], state, 0, torch_ops.F64)


def test_advance_op_and_substreams(gpu, oracle_mod):
    import torch

    from mrg8_amd import torch_ops

    rng = np.random.default_rng(8)
    n = 37
    st = rng.integers(1, 2**31 - 1, size=(n, 8)).astype(np.int32)
    offs = rng.integers(0, 2**63, size=(n, 3), dtype=np.int64)
    offs[:, 2] &= 0xFF  # offsets below 2^136
    states = torch.from_numpy(st).to("cuda:0")
    offsets = torch.from_numpy(offs).to("cuda:0")
    side = torch.cuda.Stream()
    say("torch-advance-op")
    with torch.cuda.stream(side):
        torch.ops.mrg8.advance_(states, offsets)
    side.synchronize()
    got = states.cpu().numpy().view(np.uint32)
    for i in range(n):
        x = sum(int(np.uint64(offs[i, w])) << (64 * w) for w in range(3))
        want = oracle_mod.jump_bits(st[i].view(np.uint32), [b for b in range(x.bit_length()) if (x >> b) & 1])
        assert got[i].tolist() == want.tolist(), i
    sub = torch_ops.substream_states(13579, 16, 100, device="cuda:0").cpu().numpy().view(np.uint32)
    s0 = oracle_mod.seed_state(13579)
    for i in range(16):
        want = oracle_mod.jump_bits(s0, [100 + b for b in range(i.bit_length()) if (i >> b) & 1])
        assert sub[i].tolist() == want.tolist(), i
    assert torch_ops.substream_states(1, 0, device="cuda:0").shape == (0, 8)
