"""Pins the oracle (oracle/mrg8_oracle.c) against the golden vectors of the
compiled reference, against closed forms and against independent big-integer
brute force, and shows that plausible mistakes fail these checks."""
import hashlib
import json
import os

import numpy as np
import pytest

P = 2**31 - 1
A = [1089656042, 1906537547, 1764115693, 1304127872, 189748160, 1984088114, 626062218, 1927846343]
RNORM = 1.0 / 2147483647.0
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
SEED = GOLD["seed"]


def mod1p(s: int) -> int:
    """Value of the reference's dot product for an exact integer sum s >= 0."""
    return 0 if s == 0 else 1 + (s - 1) % P


def py_seed(seed: int):
    x = seed if seed else 97531
    out = []
    for _ in range(8):
        x = (x * 6364136223846793005) % 2**64
        out.append(x >> 33)
    return out


def py_randint(state, n):
    """Independent big-integer recurrence S[n] = sum a_k S[n-1-k]."""
    st = list(state)
    out = []
    for _ in range(n):
        y = mod1p(sum(a * s for a, s in zip(A, st)))
        st = [y] + st[:7]
        out.append(y)
    return out, st


def py_matmul(X, Y):
    return [[mod1p(sum(X[r][q] * Y[q][c] for q in range(8))) for c in range(8)] for r in range(8)]


def companion():
    M = [[0] * 8 for _ in range(8)]
    M[0] = list(A)
    for r in range(1, 8):
        M[r][r - 1] = 1
    return M


def as_matrix(tab, j):
    """Matrix j of the column-major table."""
    m = tab[64 * j:64 * (j + 1)]
    return [[int(m[r + 8 * c]) for c in range(8)] for r in range(8)]


# ---- bigdot ---------------------------------------------------------------------


def test_bigdot_closed_form(oracle_mod):
    rng = np.random.default_rng(1)
    extremes = [0, 1, 2, P - 1, P, 2**31 - 1]
    for t in range(3000):
        if t < 500:
            x = rng.choice(extremes, size=8)
            y = rng.choice(extremes, size=8)
        else:
            x = rng.integers(0, 2**31, size=8)
            y = rng.integers(0, 2**31, size=8)
        s = sum(int(a) * int(b) for a, b in zip(x, y))
        assert oracle_mod.bigdot(x, y) == mod1p(s), (x, y)


def test_bigdot_never_zero_for_positive_sums(oracle_mod):
    # residue 0 with a non-zero sum is written as p ("0 is represented by 2^31-1")
    assert oracle_mod.bigdot([P] + [0] * 7, [1] * 8) == P
    assert oracle_mod.bigdot([2] + [0] * 7, [(P + 1) // 2] + [0] * 7) == 1
    assert oracle_mod.bigdot([0] * 8, [5] * 8) == 0


# ---- seeding ----------------------------------------------------------------------


def test_seed_closed_form_and_golden(oracle_mod):
    for s, want in GOLD["seed_states"].items():
        seed = int(s) & 0xFFFFFFFF  # the reference takes uint32_t
        got = oracle_mod.seed_state(seed).tolist()
        assert got == want, s
        assert py_seed(seed) == want, s
    # closed form: word k = (seed * ia^(k+1) mod 2^64) >> 33
    ia = 6364136223846793005
    for seed in (1, 7, 13579, 2**32 - 1):
        assert py_seed(seed) == [(seed * pow(ia, k + 1, 2**64)) % 2**64 >> 33 for k in range(8)]


# ---- recurrence ------------------------------------------------------------------------


def test_randint_matches_golden_and_bruteforce(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    out, st = oracle_mod.generate(s0, 1024, oracle_mod.FMT_U32)
    assert out.tolist() == [int(x) for x in GOLD["randint_first"]["first"]]
    bf, bst = py_randint(s0.tolist(), 1024)
    assert out.tolist() == bf
    assert st.tolist() == bst == GOLD["randint_first"]["state"]
    assert int(GOLD["randint_first"]["fnv1a64"], 16) == oracle_mod.fnv1a(out)


def test_states_after_n(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    for n, want in GOLD["states_after"].items():
        _, st = oracle_mod.generate(s0, int(n), oracle_mod.FMT_U32)
        assert st.tolist() == want, n


def test_bruteforce_random_states(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(20):
        st = rng.integers(0, 2**31, size=8)
        st[rng.integers(0, 8, size=2)] = rng.choice([0, P])
        got, gst = oracle_mod.generate(st, 300, oracle_mod.FMT_U32)
        bf, bst = py_randint([int(x) for x in st], 300)
        assert got.tolist() == bf and gst.tolist() == bst


def test_rand_conversions(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    u, _ = oracle_mod.generate(s0, 256, oracle_mod.FMT_U32)
    d, _ = oracle_mod.generate(s0, 256, oracle_mod.FMT_F64)
    d01, _ = oracle_mod.generate(s0, 256, oracle_mod.FMT_F64_01)
    # Python floats are IEEE doubles: one rounding of x * (1/p)
    assert d.tolist() == [float(x) * RNORM for x in u.tolist()]
    assert d01.tolist() == [float(x - 1) * RNORM for x in u.tolist()]
    assert [f"{b:016x}" for b in d.view(np.uint64)] == GOLD["rand_first"]["first"]
    assert [f"{b:016x}" for b in d01.view(np.uint64)] == GOLD["rand01_first_knl"]["first"]


def test_reciprocal_not_division():
    """x * (1/p) is what the reference computes; x / p differs (first at 4194305)."""
    assert 4194305 * RNORM != 4194305 / P
    assert P * RNORM == 1.0


def test_sequence_hashes_2_24(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    h, st = oracle_mod.hash_stream(s0, 1 << 24, oracle_mod.FMT_F64)
    assert f"{h:016x}" == GOLD["hash"]["rand_2^24"]["fnv1a64"]
    assert st.tolist() == GOLD["hash"]["rand_2^24"]["state"]
    h, _ = oracle_mod.hash_stream(s0, 1 << 24, oracle_mod.FMT_F64_01)
    assert f"{h:016x}" == GOLD["hash"]["rand01_2^24_knl"]["fnv1a64"]


@pytest.mark.parametrize("key,fmt", [("randint_2^28", 2), ("rand_2^28", 0)])
def test_sequence_hashes_2_28(oracle_mod, key, fmt):
    h, st = oracle_mod.hash_stream(oracle_mod.seed_state(SEED), 1 << 28, fmt)
    assert f"{h:016x}" == GOLD["hash"][key]["fnv1a64"]
    assert st.tolist() == GOLD["hash"][key]["state"]


def test_mrg8dnz2(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    for n, want in GOLD["mrg8dnz2"].items():
        out, st = oracle_mod.mrg8dnz2(s0, int(n))
        assert f"{oracle_mod.fnv1a(out):016x}" == want["fnv1a64"], n
        assert st.tolist() == want["state"]
        ref, rst = oracle_mod.generate(s0, int(n), oracle_mod.FMT_F64)  # same sequence as rand()
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)) and np.array_equal(st, rst)
    for n, want in GOLD["mrg8dnz2_knl"].items():  # knl variant: (x-1)/p
        ref, rst = oracle_mod.generate(s0, int(n), oracle_mod.FMT_F64_01)
        assert f"{oracle_mod.fnv1a(ref):016x}" == want["fnv1a64"], n


# ---- jump table --------------------------------------------------------------------


def test_jump_table_bytes_match_reference_file(oracle_mod, tmp_path):
    p = tmp_path / "jump_matrix.txt"
    assert oracle_mod.lib().mrg8o_write_jump_table(
        oracle_mod.jump_table().ctypes.data_as(oracle_mod.ctypes.POINTER(oracle_mod.ctypes.c_uint32)),
        str(p).encode()) == 0
    assert hashlib.md5(p.read_bytes()).hexdigest() == GOLD["jump_matrix_md5"]


def test_jump_table_closed_form(oracle_mod):
    tab = oracle_mod.jump_table()
    assert as_matrix(tab, 0) == companion()
    for j in (0, 1, 2, 5, 30, 100, 245):
        M = as_matrix(tab, j)
        assert as_matrix(tab, j + 1) == py_matmul(M, M), j
    # A^8 is entry 3
    A8 = companion()
    for _ in range(3):
        A8 = py_matmul(A8, A8)
    assert as_matrix(tab, 3) == A8


# ---- jumps ----------------------------------------------------------------------


def test_jumps_match_golden(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    for j in GOLD["jumps"]:
        step = int(j["step"])
        assert oracle_mod.jump(s0, step).tolist() == j["state"], step
        assert oracle_mod.jump_vec(s0, step).tolist() == j["state"], step
    for j in GOLD["jump_bits"]:
        assert oracle_mod.jump_bits(s0, j["bits"]).tolist() == j["state"], j["bits"]


def test_jump_invariants(oracle_mod):
    rng = np.random.default_rng(5)
    s0 = oracle_mod.seed_state(SEED)
    assert np.array_equal(oracle_mod.jump(s0, 0), s0)
    for n in (1, 2, 7, 8, 9, 64, 1000, 99991):
        _, st = oracle_mod.generate(s0, n, oracle_mod.FMT_U32)
        assert np.array_equal(oracle_mod.jump(s0, n), st), n
    for _ in range(50):
        a = int(rng.integers(0, 2**62, dtype=np.uint64))
        b = int(rng.integers(0, 2**62, dtype=np.uint64))
        assert np.array_equal(oracle_mod.jump(oracle_mod.jump(s0, a), b), oracle_mod.jump(s0, a + b))
    # 200-bit jumps compose with 64-bit ones: 2^64 = bit 64
    x = oracle_mod.jump_bits(s0, [64])
    assert np.array_equal(oracle_mod.jump(oracle_mod.jump(x, 2**63), 2**63), oracle_mod.jump_bits(s0, [65]))


def test_matrix_chain_equals_matrix_product_with_zero_words(oracle_mod):
    """The matrix-vector chain (knl form) equals the matrix-product form
    (original form) bit for bit, also for states with 0 and p words."""
    rng = np.random.default_rng(11)
    for t in range(400):
        st = rng.integers(0, 2**31, size=8).astype(np.uint32)
        st[rng.integers(0, 8, size=int(rng.integers(0, 8)))] = rng.choice([0, P])
        step = int(rng.integers(0, 40)) if t % 2 else int(rng.integers(0, 2**64, dtype=np.uint64))
        assert np.array_equal(oracle_mod.jump(st, step), oracle_mod.jump_vec(st, step)), (st, step)


def test_jump_then_randint_golden(oracle_mod):
    s0 = oracle_mod.seed_state(SEED)
    for r in GOLD["jump_then_randint"]:
        out, st = oracle_mod.generate(oracle_mod.jump(s0, int(r["step"])), 256, oracle_mod.FMT_U32)
        assert f"{oracle_mod.fnv1a(out):016x}" == r["fnv1a64"]
        assert st.tolist() == r["state"]


def test_set_state_streams_golden(oracle_mod):
    for r in GOLD["set_state_streams"]:
        out, st = oracle_mod.generate(np.array(r["init"], dtype=np.uint32), r["n"], r["fmt"])
        assert f"{oracle_mod.fnv1a(out):016x}" == r["fnv1a64"]
        assert st.tolist() == r["state"]


def test_noncanonical_golden(oracle_mod):
    for r in GOLD["noncanonical"]:
        w = np.array(r["init"], dtype=np.uint32)
        out, st = oracle_mod.generate(oracle_mod.jump(w, int(r["step"])), 300, oracle_mod.FMT_U32)
        assert f"{oracle_mod.fnv1a(out):016x}" == r["fnv1a64"], (r["init"], r["step"])
        assert [str(x) for x in out[:24]] == r["first"]
        assert st.tolist() == r["state"]


# ---- the checks catch plausible mistakes ------------------------------------------------


def fold_input_above_p():
    """Inputs whose sum survives two Mersenne folds above p, so that only the
    third fold brings the result into [1, p]."""
    T = 2**63 + 2**31 - 1  # first fold -> 3*2^31 - 1, second fold -> 2^31 + 1
    x = [P, P, P, 0, 0, 0, 0, 0]
    y = [P, P, 0, 1, 0, 0, 0, 0]
    rem = T - 2 * P * P
    y[2] = rem // P
    x[3] = rem - P * y[2]
    assert sum(a * b for a, b in zip(x, y)) == T and max(x + y) < 2**31
    return x, y


def two_fold_dot(x, y):
    s1 = sum(a * b for a, b in zip(x[:4], y[:4]))
    s2 = sum(a * b for a, b in zip(x[4:], y[4:]))
    s = (s1 & P) + (s1 >> 31) + (s2 & P) + (s2 >> 31)
    return (s & P) + (s >> 31)


def test_bigdot_third_fold(oracle_mod):
    x, y = fold_input_above_p()
    want = mod1p(sum(a * b for a, b in zip(x, y)))
    assert oracle_mod.bigdot(x, y) == want
    assert two_fold_dot(x, y) > P and two_fold_dot(x, y) != want  # a dropped fold is caught here


def test_mutations_are_caught(oracle_mod):
    """Plausible mistakes, planted in Python copies of the algorithm, each
    disagree with the golden vectors that pin the oracle."""
    import ctypes

    s0 = oracle_mod.seed_state(SEED).tolist()
    gold = [int(x) for x in GOLD["randint_first"]["first"]]
    # coefficients reversed
    out, st = [], list(s0)
    for _ in range(16):
        y = mod1p(sum(a * s for a, s in zip(A[::-1], st)))
        st = [y] + st[:7]
        out.append(y)
    assert out != gold[:16]
    # state kept oldest-first
    out, st = [], list(s0)[::-1]
    for _ in range(16):
        y = mod1p(sum(a * s for a, s in zip(A, st)))
        st = st[1:] + [y]
        out.append(y)
    assert out != gold[:16]
    # transposed jump matrices
    tab = oracle_mod.jump_table().copy()
    for j in range(64):
        tab[64 * j:64 * (j + 1)] = tab[64 * j:64 * (j + 1)].reshape(8, 8).T.reshape(-1)
    bad = np.zeros(8, dtype=np.uint32)
    st = np.array(s0, dtype=np.uint32)
    p32 = ctypes.POINTER(ctypes.c_uint32)
    oracle_mod.lib().mrg8o_jump_vec(tab.ctypes.data_as(p32), ctypes.c_uint64(123456789), st.ctypes.data_as(p32),
                                    bad.ctypes.data_as(p32))
    gold_jump = {j["step"]: j["state"] for j in GOLD["jumps"]}
    step = next(k for k in gold_jump if int(k) > 1000)
    assert oracle_mod.jump(st, int(step)).tolist() == gold_jump[step]
    assert bad.tolist() != oracle_mod.jump(st, 123456789).tolist()
    # division instead of the reciprocal multiply
    d_gold = [int(b, 16) for b in GOLD["rand_first"]["first"]]
    u = gold[:256]
    by_div = [int(np.float64(float(x) / P).view(np.uint64)) for x in u]
    by_mul = [int(np.float64(float(x) * RNORM).view(np.uint64)) for x in u]
    assert by_mul == d_gold
    big, _ = oracle_mod.generate(np.array(s0, dtype=np.uint32), 1 << 16, oracle_mod.FMT_U32)
    assert any(float(x) / P != float(x) * RNORM for x in big.tolist())
    assert by_div != d_gold  # 3 of these 256 golden doubles differ under division
    # wrong seed multiplier
    x, wrong = SEED, []
    for _ in range(8):
        x = (x * 6364136223846793007) % 2**64
        wrong.append(x >> 33)
    assert wrong != GOLD["seed_states"][str(SEED)]


def test_wide_jump_matrices_have_no_zero_entries(oracle_mod):
    """A^(d 256^k) for byte positions k = 8..31 (the device's table for
    offsets past 64 bits) has no entry equal to 0 mod p, so the exact-zero
    case of bigDotProd (original/mrg8.cpp:72-82) depends only on the state's
    zero words there, as the device's support masks assume."""
    p = 2**31 - 1
    c = [int(x) for x in oracle_mod.coeffs()]
    a = [c] + [[1 if j == r - 1 else 0 for j in range(8)] for r in range(1, 8)]

    def mul(x, y):
        return [[sum(x[r][k] * y[k][j] for k in range(8)) % p for j in range(8)] for r in range(8)]

    m = a
    for _ in range(64):
        m = mul(m, m)  # A^(2^64)
    for _k in range(8, 32):
        q = m
        for d in range(1, 256):
            if d > 1:
                q = mul(q, m)
            assert all(v != 0 for row in q for v in row)
        for _ in range(8):
            m = mul(m, m)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1000, 65537, (1 << 20) + 3])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_vector_baseline_equals_serial(oracle_mod, n, threads):
    """The vector + OpenMP restatement of knl/mrg8_vec.h (the timed CPU
    baseline) gives the serial generator's [0,1) doubles bit for bit: the
    reference's only bit-exact check, knl/tests/test_mrg8_vec_outer.cpp:70-83
    (seed 13579, mrg8_vec_outer against mrg8::rand), asserted here."""
    st = oracle_mod.seed_state(13579)
    got, _, used = oracle_mod.vec_outer_tp(st, n, threads)
    want, _ = oracle_mod.generate(st, n, 1)
    assert used >= 1
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


# ---- the checker's checksums -------------------------------------------------
# oracle.wsum / oracle.stream_wsums decide the full-size batched and served
# checks; they are pinned here by closed forms and against an independent
# big-integer evaluation of their definition sum_i w_i (2i+1) mod 2^64.

def _wsum_bigint(arr):
    w = arr.view(np.uint64) if arr.dtype == np.float64 else arr.astype(np.uint64)
    return sum(int(x) * (2 * i + 1) for i, x in enumerate(w.tolist())) % (1 << 64)


def test_wsum_closed_forms(oracle_mod):
    from conftest import wsum as np_wsum

    for n in (0, 1, 2, 7, 1000, 4097):
        ones = np.ones(n, dtype=np.uint32)
        assert oracle_mod.wsum(ones) == n * n  # sum of the first n odd numbers
        assert np_wsum(ones) == n * n
    # one non-zero word at position i weighs 2i + 1; f64 enters by its bit pattern
    z = np.zeros(50, dtype=np.uint32)
    z[17] = 0xFFFFFFFF
    assert oracle_mod.wsum(z) == 0xFFFFFFFF * 35
    d = np.zeros(9, dtype=np.float64)
    d[4] = 1.0
    assert oracle_mod.wsum(d) == (0x3FF0000000000000 * 9) % (1 << 64)
    d[4] = -2.0  # the sign bit counts: 0xC000000000000000 * 9 wraps mod 2^64
    assert oracle_mod.wsum(d) == (0xC000000000000000 * 9) % (1 << 64)


def test_wsum_equals_definition(oracle_mod):
    from conftest import wsum as np_wsum

    s0 = oracle_mod.seed_state(4242)
    for fmt in (0, 1, 2):
        for n in (1, 63, 4096, 4099):
            out, _ = oracle_mod.generate(s0, n, fmt)
            want = _wsum_bigint(out)
            assert oracle_mod.wsum(out) == want, (fmt, n)
            assert np_wsum(out) == want, (fmt, n)


def test_stream_wsums_and_windows_equal_generated(oracle_mod):
    """Per-stream checksums and windows at non-zero offsets equal those of the
    serial generator jumped there (counts across the 4096-output block edge)."""
    rng = np.random.default_rng(99)
    k = 12
    states = rng.integers(1, 2**31 - 1, size=(k, 8), dtype=np.int64).astype(np.uint32)
    offsets = rng.integers(0, 2**40, size=k, dtype=np.uint64)
    offsets[0] = 0
    counts = np.array([1, 64, 4095, 4096, 4097, 9000, 3, 5000, 8192, 100, 257, 4100], dtype=np.uint64)
    fmts = np.array([0, 1, 2] * 4, dtype=np.int32)
    sums = oracle_mod.stream_wsums(states, offsets, counts, fmts)
    for i in range(k):
        st = oracle_mod.jump(states[i], int(offsets[i]))
        out, _ = oracle_mod.generate(st, int(counts[i]), int(fmts[i]))
        assert int(sums[i]) == _wsum_bigint(out), i
    for fmt in (0, 1, 2):
        win = oracle_mod.stream_windows(states, offsets, 40, fmt)
        for i in range(k):
            out, _ = oracle_mod.generate(oracle_mod.jump(states[i], int(offsets[i])), 40, fmt)
            assert win[i].tobytes() == out.tobytes(), (fmt, i)
