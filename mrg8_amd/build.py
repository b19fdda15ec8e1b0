"""Builds libmrg8_amd.so (gfx950 kernels + C ABI) and libmrg8_probe.so (the
benchmark's measurement kernels) in-tree with hipcc.

    python -m mrg8_amd.build            # compile (always rebuilds)
    python -m mrg8_amd.build --check    # print the command lines only
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libmrg8_amd.so")
PROBE_SRC = os.path.join(CSRC, "probe", "mrg8_probe.hip")
PROBE_OUT = os.path.join(HERE, "libmrg8_probe.so")
OBJDIR = os.path.join(ROOT, "_build", "mrg8_amd")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-I" + os.path.join(ROOT, "include")]
# Device code: gfx950 only; packed FP32 instructions are disabled (they are
# not needed here and some forms have proven unreliable on this hardware).
# MFMA results go to ordinary VGPRs: the matrix-core fill reads every result
# with VALU instructions, and results left in AGPRs cost one v_accvgpr_read
# per value (four per output) -- 189 VGPRs either way fit its two waves per SIMD.
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def compile_cmds():
    cc = hipcc()
    cmds = []
    for src in sources():
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        flags = COMMON + (DEVICE if src.endswith(".hip") else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"])
        cmds.append((obj, [cc] + flags + ["-c", src, "-o", obj]))
    return cmds


def link_cmd(objs):
    return [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", OUT + ".tmp"] + objs


def probe_cmd():
    return [hipcc()] + COMMON + DEVICE + ["-shared", "-o", PROBE_OUT + ".tmp", PROBE_SRC]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    cmds = compile_cmds()

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        noise = ("packed-fp32-ops", "unknown target feature", "--hip-link")
        err = "\n".join(l for l in r.stderr.splitlines() if not any(n in l for n in noise))
        if r.returncode != 0:
            raise RuntimeError("compile failed:\n" + " ".join(cmd) + "\n" + r.stderr)
        return err

    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds) + 1)) as ex:
        probe = ex.submit(run, probe_cmd())
        for err in ex.map(run, [c for _, c in cmds]):
            if err.strip() and verbose:
                print(err, file=sys.stderr)
        probe.result()
    run(link_cmd([o for o, _ in cmds]))
    os.replace(OUT + ".tmp", OUT)
    os.replace(PROBE_OUT + ".tmp", PROBE_OUT)
    return OUT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    if a.check:
        for _, c in compile_cmds():
            print(" ".join(c))
        print(" ".join(link_cmd(["<objs>"])))
        print(" ".join(probe_cmd()))
        return
    print(build(verbose=a.verbose))


if __name__ == "__main__":
    main()
