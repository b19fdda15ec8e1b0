"""The reference's benchmark drivers, on the C ABI.

Counterpart of cuda/tests/test_mrg8_outer.cu and cuda/tests/test_mrg8_inner.cu
with the conventions of common/rng_test.h: seed 13579, N = argv * 2^20
doubles, a sweep of total threads TNUM = 1024 .. 2^20 and threads per block
BS = 32 .. 1024 (BS <= TNUM), ITER = 10 timed calls per point of which the
first is dropped and the other 9 averaged, check_rand's mean and standard
deviation printed for the first call of the sweep (the reference prints them
for every point), and one line per point (the BS field empty for inner):

    EVALUATION, <name>, <TNUM>, <BS>, <N>, <million rng/s>, <ms>

Each call is the blocking drop-in (mrg8_outer_32 / mrg8_inner: [0, 1)
doubles, seed_init before every call as the reference does).  TNUM and BS
only choose the launch shape here (BS below 64 launches 64-thread blocks:
one wave); the outputs are identical for every shape.  `--auto` adds a line
for the library's own shape (mrg8_generate).

    python -m mrg8_amd.harness 256            # N = 256 * 2^20
    python -m mrg8_amd.harness 256 --inner
"""
import argparse
import math
import sys

import numpy as np

ITER = 10  # common/rng_test.h:24


def check_rand(ran: np.ndarray) -> None:
    """common/rng_test.h:29-47: mean and standard deviation, printed."""
    ave = float(np.mean(ran))
    print(f"Arithmetic mean: {ave}")
    print(f"Standard deviation: {math.sqrt(float(np.mean((ran - ave) ** 2)))}")


def main(argv=None) -> int:
    import mrg8_amd
    from mrg8_amd import hip

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("n_mb", type=int, nargs="?", default=1, help="N = n_mb * 2^20 doubles (default 1)")
    ap.add_argument("--inner", action="store_true", help="time mrg8_inner (test_mrg8_inner.cu) instead")
    ap.add_argument("--tnum-max", type=int, default=1 << 20)
    ap.add_argument("--auto", action="store_true", help="also time the library's own launch shape")
    a = ap.parse_args(argv)
    n = a.n_mb * 1024 * 1024
    if n > 2**31 - 1:
        print("N must fit the reference's int", file=sys.stderr)
        return 2
    seed = 13579  # cuda/tests/test_mrg8_outer.cu:35
    name = "MRG8_HIP_inner" if a.inner else "MRG8_HIP_outer"
    print(f"Generating {n} of 64-bit floating random numbers")
    d_ran = hip.DeviceBuffer(n * 8)
    m = mrg8_amd.mrg8_cuda(seed)
    e0, e1 = hip.Event(), hip.Event()
    checked = False

    def timed(call):
        nonlocal checked
        total = 0.0
        for i in range(ITER):
            m.seed_init(seed)
            e0.record()
            call()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_ms(e1)
            if i == 0 and not checked:
                check_rand(d_ran.copy_to_host(np.float64, n))
                checked = True
            if i > 0:
                total += ms
        return total / (ITER - 1)

    tnum = 1024
    while tnum <= a.tnum_max:
        bss = [tnum] if a.inner else [bs for bs in (32, 64, 128, 256, 512, 1024) if bs <= tnum]
        for bs in bss:
            if a.inner:
                ms = timed(lambda: m.mrg8_inner(d_ran, n, tnum))
            else:
                ms = timed(lambda: m.mrg8_outer_32(d_ran, n, bs, tnum))
            mrng = n / ms / 1000.0
            bs_field = "" if a.inner else str(bs)  # the inner driver leaves BS empty
            print(f"EVALUATION, {name}, {tnum}, {bs_field}, {n}, {mrng:f}, {ms:f}", flush=True)
        tnum *= 2
    if a.auto:
        ms = timed(lambda: (m.generate(d_ran, n, mrg8_amd.F64_01), m.synchronize()))
        print(f"EVALUATION, {name}_auto, 0, 0, {n}, {n / ms / 1000.0:f}, {ms:f}", flush=True)
    d_ran.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
