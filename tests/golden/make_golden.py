"""Writes tests/golden/ref_vectors.json from the reference generator compiled
from its own sources (oracle/_ref, built by `make -C oracle ref` where the
reference tree is present).  The fixture holds inputs and outputs only.

    python tests/golden/make_golden.py

Builds run: ref_original (original/mrg8.cpp), ref_cuda_host (cuda/mrg8.cpp),
ref_knl (knl/mrg8.h); see oracle/ref_driver.cpp for the command protocol.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
REF_SRC = os.environ.get("MRG8_REF", "/root/reference")
SEED = 13579
MASK64 = (1 << 64) - 1


def splitmix64(x: int):
    """Fixed-seed source of test inputs (steps, states)."""
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


class Ref:
    def __init__(self, name: str):
        exe = os.path.join(REF, name)
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: run `make -C oracle ref` first")
        self.p = subprocess.Popen([exe], cwd=REF, stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    def ask(self, line: str) -> str:
        self.p.stdin.write(line + "\n")
        self.p.stdin.flush()
        return self.p.stdout.readline().strip()

    def close(self):
        self.p.stdin.write("Q\n")
        self.p.stdin.flush()
        self.p.wait()


def md5_of(path: str) -> str:
    import hashlib

    with open(path, "rb") as f:
        return hashlib.md5(f.read()).hexdigest()


def parse_gen(ans: str):
    head, state = ans.split("|")
    parts = head.split()
    return {"fnv1a64": parts[0], "first": parts[1:], "state": [int(x) for x in state.split()]}


def gen(ref: Ref, seed, state, step, n, fmt, k):
    s = "-" if seed is None else str(seed)
    w = " ".join(str(x) for x in (state or [0] * 8))
    return parse_gen(ref.ask(f"G {s} {w} {step} {n} {fmt} {k}"))


def main():
    orig, cuda, knl = Ref("ref_original"), Ref("ref_cuda_host"), Ref("ref_knl")
    rnd = splitmix64(20240601)
    fx = {"about": "outputs of the reference compiled from original/, cuda/ (host class) and knl/ sources; "
                   "fnv1a64 is over the little-endian bytes of the outputs (u32 words or IEEE doubles)",
          "seed": SEED,
          # md5 of the reference's jump_matrix.txt (identical in ./, original/, cuda/, knl/)
          "jump_matrix_md5": md5_of(os.path.join(REF_SRC, "jump_matrix.txt"))}

    seeds = [13579, 0, 1, 2, 97531, 2147483647, 4294967295, 123456789, 1 << 31]
    fx["seed_states"] = {str(s): [int(x) for x in orig.ask(f"S {s}").split()] for s in seeds}

    fx["randint_first"] = gen(orig, SEED, None, 0, 1024, 2, 1024)
    fx["rand_first"] = gen(orig, SEED, None, 0, 256, 0, 256)
    fx["rand01_first_knl"] = gen(knl, SEED, None, 0, 256, 0, 256)
    fx["states_after"] = {str(n): gen(orig, SEED, None, 0, n, 2, 0)["state"]
                          for n in (1, 7, 8, 9, 1000, 1 << 20, 1 << 24)}

    # full-sequence hashes of the benchmark shapes
    fx["hash"] = {
        "rand_2^24": gen(orig, SEED, None, 0, 1 << 24, 0, 4),
        "rand01_2^24_knl": gen(knl, SEED, None, 0, 1 << 24, 0, 4),
        "randint_2^28": gen(orig, SEED, None, 0, 1 << 28, 2, 4),
        "rand_2^28": gen(orig, SEED, None, 0, 1 << 28, 0, 4),
    }

    steps = [0, 1, 2, 3, 7, 8, 9, 15, 16, 17, 1 << 31, (1 << 32) + 5, (1 << 63) - 1, MASK64]
    steps += [next(rnd) >> (next(rnd) % 64) for _ in range(32)]
    jumps = []
    for st in steps:
        o = gen(orig, SEED, None, st, 0, 2, 0)["state"]
        c = gen(cuda, SEED, None, st, 0, 2, 0)["state"]
        k = gen(knl, SEED, None, st, 0, 2, 0)["state"]
        assert o == c == k, (st, o, c, k)
        jumps.append({"step": str(st), "state": o})
    fx["jumps"] = jumps

    bitsets = [[100], [150, 3], [199, 64, 65, 0], [63, 64], list(range(0, 200, 7))]
    bitsets += [sorted({next(rnd) % 200 for _ in range(12)}) for _ in range(6)]
    fx["jump_bits"] = []
    for bs in bitsets:
        arg = ",".join(str(b) for b in bs)
        o = [int(x) for x in orig.ask(f"B {SEED} {arg}").split()]
        c = [int(x) for x in cuda.ask(f"B {SEED} {arg}").split()]
        assert o == c, (bs, o, c)
        fx["jump_bits"].append({"bits": bs, "state": o})

    fx["mrg8dnz2"] = {str(n): gen(cuda, SEED, None, 0, n, 3, 8) for n in (1, 1023, 1024, 1025, 5000)}
    fx["mrg8dnz2_knl"] = {str(n): gen(knl, SEED, None, 0, n, 3, 8) for n in (1, 1025, 5000)}

    # jump-heavy shape: random 63-bit steps from one seed, then 256 words each
    fx["jump_then_randint"] = []
    for _ in range(64):
        st = next(rnd) >> 1
        r = gen(orig, SEED, None, st, 256, 2, 4)
        fx["jump_then_randint"].append({"step": str(st), **r})

    # set_state streams of uneven length (batched shape, small)
    fx["set_state_streams"] = []
    for i in range(24):
        w = [1 + next(rnd) % (2**31 - 2) for _ in range(8)]
        n = 1 + next(rnd) % 200000
        fmt = (0, 2)[i % 2]
        r = gen(orig, None, w, 0, n, fmt, 4)
        fx["set_state_streams"].append({"init": w, "n": n, "fmt": fmt, **r})

    # states with zero / p words (non-canonical), generated through set_state
    fx["noncanonical"] = []
    for w in ([0, 0, 0, 0, 0, 0, 0, 5], [0, 2147483647, 0, 0, 0, 0, 0, 0], [2147483647] * 8, [0] * 8,
              [1, 0, 1, 0, 1, 0, 1, 0]):
        for step in (0, 1, 3, 9, 100, 123456789):
            r = gen(orig, None, w, step, 300, 2, 24)
            fx["noncanonical"].append({"init": w, "step": str(step), **r})

    for r in (orig, cuda, knl):
        r.close()
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=0, separators=(",", ":"))
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
