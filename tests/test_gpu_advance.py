"""GPU parity of the bulk device jump (mrg8_advance_states) against the oracle.

Offsets below 2^200 are checked against the oracle's 200-bit jump
(original/mrg8.cpp:105-122, pinned by the golden jump vectors); offsets up to
2^256 against the closed form: A^x mod p by square-and-multiply over Python
integers, applied with bigDotProd's semantics (original/mrg8.cpp:72-82), the
closed form itself checked against the oracle below 2^200 in the same test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2**31 - 1


def say(*a):
    print("[case]", *a, flush=True)


def split_words(x: int, words: int):
    return [(x >> (64 * i)) & (2**64 - 1) for i in range(words)]


def set_bits(x: int):
    return [i for i in range(x.bit_length()) if (x >> i) & 1]


def run_advance(gpu, states: np.ndarray, offsets, words: int, in_place: bool = False) -> np.ndarray:
    from mrg8_amd.hip import DeviceBuffer

    n = len(states)
    d_in = DeviceBuffer(max(1, n) * 32)
    d_in.copy_from_host(np.ascontiguousarray(states, dtype=np.uint32))
    offs = np.array([w for x in offsets for w in split_words(int(x), words)], dtype=np.uint64)
    d_off = DeviceBuffer(max(1, offs.size) * 8)
    if offs.size:
        d_off.copy_from_host(offs)
    d_out = d_in if in_place else DeviceBuffer(max(1, n) * 32)
    gpu.advance_states(d_in, d_off, words, d_out, n)
    gpu.hip.synchronize()
    got = d_out.copy_to_host(np.uint32, n * 8).reshape(n, 8)
    d_in.free()
    d_off.free()
    if not in_place:
        d_out.free()
    return got


def closed_form_jump(coeffs, state, x: int) -> list:
    """state advanced by x: A^x (A the companion matrix of the recurrence,
    state newest word first) by square-and-multiply mod p, each new word a
    bigDotProd of a row of A^x with the state."""
    a = [[int(c) for c in coeffs]] + [[1 if c == r - 1 else 0 for c in range(8)] for r in range(1, 8)]

    def mul(x_, y_):
        return [[sum(x_[r][k] * y_[k][c] for k in range(8)) % P for c in range(8)] for r in range(8)]

    m = [[1 if r == c else 0 for c in range(8)] for r in range(8)]
    base = a
    while x:
        if x & 1:
            m = mul(m, base)
        base = mul(base, base)
        x >>= 1
    out = []
    for r in range(8):
        s = sum(m[r][c] * int(state[c]) for c in range(8))
        out.append(0 if s == 0 else 1 + (s - 1) % P)
    return out


def test_advance_64bit_offsets(gpu, oracle_mod):
    rng = np.random.default_rng(71)
    n = 300
    states = rng.integers(1, P + 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    offs = [int(x) for x in rng.integers(0, 2**63, size=n, dtype=np.uint64)]
    offs[:4] = [0, 1, 2**64 - 1, 256]
    say("advance words=1", n)
    got = run_advance(gpu, states, offs, 1)
    for i in range(n):
        assert got[i].tolist() == oracle_mod.jump(states[i], offs[i]).tolist(), i


@pytest.mark.parametrize("words", [2, 3, 4])
def test_advance_wide_offsets_below_2e200(gpu, oracle_mod, words):
    rng = np.random.default_rng(72 + words)
    n = 61  # not a multiple of the 8 states a wave takes
    states = rng.integers(1, P + 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    top = min(64 * words, 200)
    offs = [int.from_bytes(rng.bytes(32), "little") % (1 << top) for _ in range(n)]
    offs[0] = 0
    offs[1] = (1 << (top - 1))  # the highest single bit
    offs[2] = (1 << 64) + 5  # a carry into the second word
    say("advance words", words, n)
    got = run_advance(gpu, states, offs, words)
    for i in range(n):
        want = oracle_mod.jump_bits(states[i], set_bits(offs[i]))
        assert got[i].tolist() == want.tolist(), (words, i, hex(offs[i]))


def test_advance_full_256bit_offsets_closed_form(gpu, oracle_mod):
    rng = np.random.default_rng(99)
    coeffs = oracle_mod.coeffs()
    # the closed form agrees with the oracle's 200-bit jump
    s0 = oracle_mod.seed_state(13579)
    for x in (1, 12345, (1 << 150) + 7, (1 << 199) + (1 << 64) + 3):
        assert closed_form_jump(coeffs, s0, x) == oracle_mod.jump_bits(s0, set_bits(x)).tolist(), hex(x)
    n = 6
    states = rng.integers(1, P + 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    offs = [int.from_bytes(rng.bytes(32), "little") for _ in range(n)]
    offs[0] = 2**256 - 1
    offs[1] = 1 << 255
    offs[2] = (1 << 248) + (1 << 200)  # byte positions past the 247 jump_matrix.txt entries
    say("advance words=4 full width", n)
    got = run_advance(gpu, states, offs, 4)
    for i in range(n):
        assert got[i].tolist() == closed_form_jump(coeffs, states[i], offs[i]), (i, hex(offs[i]))


def test_advance_zero_words_and_in_place(gpu, oracle_mod):
    rng = np.random.default_rng(5)
    n = 40
    states = rng.integers(1, P + 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    for i in range(0, n, 3):
        states[i, rng.integers(0, 8, size=int(rng.integers(1, 7)))] = 0  # some zero words
    states[5] = 0  # the all-zero state stays zero
    states[6] = P  # words equal to p (the reference's [1, p] representation)
    offs = [int.from_bytes(rng.bytes(16), "little") for _ in range(n)]
    offs[7] = 3  # small offsets multiply by rows of A^d with structural zeros
    say("advance zero words, in place")
    got = run_advance(gpu, states, offs, 2, in_place=True)
    for i in range(n):
        want = oracle_mod.jump_bits(states[i], set_bits(offs[i]))
        assert got[i].tolist() == want.tolist(), (i, states[i].tolist(), hex(offs[i]))


def test_advance_feeds_a_fill(gpu, oracle_mod):
    """States advanced on the device by 2^100 start fills whose outputs match
    the oracle's generation from the oracle's jump."""
    from mrg8_amd.hip import DeviceBuffer

    s0 = oracle_mod.seed_state(13579)
    st = run_advance(gpu, s0.reshape(1, 8), [1 << 100], 2)[0]
    assert st.tolist() == oracle_mod.jump_bits(s0, [100]).tolist()
    g = gpu.mrg8_cuda(1)
    g.set_state(st)
    n = 100_000
    buf = DeviceBuffer(n * 8)
    g.generate(buf, n, 0)
    g.synchronize()
    want, _ = oracle_mod.generate(st, n, 0)
    assert np.array_equal(buf.copy_to_host(np.float64, n).view(np.uint64), want.view(np.uint64))
    buf.free()


def test_advance_argument_errors(gpu):
    from mrg8_amd.hip import DeviceBuffer

    lib = gpu.capi
    d = DeviceBuffer(64)
    with pytest.raises(lib.Mrg8Error) as e:
        gpu.advance_states(d, d, 5, d, 1)  # words above 4
    assert e.value.code == lib.ERR_ARG
    with pytest.raises(lib.Mrg8Error) as e:
        gpu.advance_states(d.ptr + 2, d, 1, d, 1)  # misaligned states
    assert e.value.code == lib.ERR_ARG
    gpu.advance_states(d, d, 1, d, 0)  # n = 0 is a no-op
    d.free()


def test_advance_p_words_on_the_centred_path(gpu, oracle_mod):
    """Waves whose eight states have no zero word take the centred fast path;
    put words equal to p (the reference's representation of residue 0) there:
    all-p states, mixes of p and ordinary words, and single p words, eight
    states per wave (8 lanes per state, 64 per wave) with no zero word."""
    rng = np.random.default_rng(17)
    n = 32  # four waves of eight states
    states = rng.integers(1, P, size=(n, 8), dtype=np.int64).astype(np.uint32)
    states[0:8] = P  # wave 0: every word p
    for i in range(8, 16):  # wave 1: half the words p
        states[i, rng.permutation(8)[:4]] = P
    for i in range(16, 24):  # wave 2: one p word each, at every position
        states[i, i - 16] = P
    assert (states != 0).all()  # wave 3 (24..31): ordinary words only
    for words, bits in ((1, 64), (4, 200)):
        offs = [int.from_bytes(rng.bytes(32), "little") & ((1 << bits) - 1) for _ in range(n)]
        offs[0], offs[9], offs[17] = 1, 255, (1 << 64) - 1 if words == 1 else (1 << 199) + 3
        say("advance p words on the centred path", words)
        got = run_advance(gpu, states, offs, words)
        for i in range(n):
            want = oracle_mod.jump_bits(states[i], set_bits(offs[i]))
            assert got[i].tolist() == want.tolist(), (words, i, states[i].tolist(), hex(offs[i]))
