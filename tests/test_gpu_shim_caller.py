"""A C++ program written against include/mrg8_cuda.hpp (the reference's class
names) runs on the device with the reference driver's call shape
(cuda/tests/test_mrg8_outer.cu:46-61) and reproduces the reference's own
[0, 1) sequence: the 2^24-double hash pinned by the compiled knl reference."""
import os
import subprocess

import pytest

from test_gpu_parity import GOLD, SEED

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "shim_caller")


def expected_print_state(st):
    # original/mrg8.cpp:41-46: setw(10) and ", " after the first seven words
    return "(S[n-1], S[n-2], ..., S[n-8])= (" + "".join(f"{int(x):>10}, " for x in st[:7]) + f"{int(st[7])})"


@pytest.mark.parametrize("bs,tnum", [(256, 65536), (64, 1024), (1024, 1 << 20)])
def test_cpp_caller_reproduces_reference_sequence(gpu, oracle_mod, bs, tnum):
    assert os.path.exists(EXE), "tests/cpp/shim_caller not built (run build())"
    n = 1 << 24
    print("[case] shim_caller", n, bs, tnum, flush=True)
    r = subprocess.run([EXE, str(n), str(bs), str(tnum)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    tags = dict(l.split(" ", 1) for l in lines if l.split(" ", 1)[0] in ("outer", "outer2", "inner", "host"))
    want01 = GOLD["hash"]["rand01_2^24_knl"]
    assert tags["outer"] == want01["fnv1a64"]
    assert tags["outer2"] == want01["fnv1a64"]
    assert tags["inner"] == want01["fnv1a64"]
    # mrg8::rand(double*, int) into host memory: original/mrg8.cpp's (0, 1] sequence
    assert tags["host"] == GOLD["hash"]["rand_2^24"]["fnv1a64"]
    # state after 2^24 outputs, printed in the reference's format
    assert expected_print_state(want01["state"]) in lines
    # print_matrix of A: row 0 the coefficients, then the shifted identity
    a = [[0] * 8 for _ in range(8)]
    a[0] = [int(x) for x in oracle_mod.coeffs()]
    for i in range(1, 8):
        a[i][i - 1] = 1
    mat = ["".join(f"{v:>10} " for v in row) for row in a]
    assert lines[-8:] == mat
