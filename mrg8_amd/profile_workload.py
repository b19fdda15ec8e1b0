"""Fixed workload for rocprofv3 runs (kernel traces and HBM counters).

    rocprofv3 --kernel-trace --stats -d OUT -o run -- python3 -m mrg8_amd.profile_workload
    rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 -m mrg8_amd.profile_workload
    rocprofv3 --pmc WRITE_SIZE -d OUT -o write -- python3 -m mrg8_amd.profile_workload

Dispatches, in order (each `--warmup` + `--reps` times; summaries use the
last `--reps` dispatches of each kind):
  1. a runtime device-to-device copy of 1 GiB  -- the calibration kernel of
     known traffic (1 GiB read + 1 GiB written)
  2. single-stream fill, 2^28 float64 outputs (seed 13579), default shape
     (the vector kernel: 512-thread blocks in XCD runs of 64)
  3. single-stream fill, 2^28 uint32 outputs, default shape (the
     matrix-core fill)
  3a. the same word fill with the vector kernel, and the float64 fill with
     the matrix-core fill (outputs_per_lane 16, 1024-thread blocks, 2048 per
     wave)
  4. batched launch: 4096 streams, log-uniform lengths, float64, one
     2048-output tile per wave, 1024-thread blocks in XCD runs of 32; then
     the same batch on a persistent grid of 512-thread blocks, 2 per CU,
     taking 2048-output tiles in address order                     (--batched)
  5. jump-heavy batched launch: 65536 x (63-bit jump + 256 uint32) (--batched)
  6. bulk advance of 2^20 device-resident states by random 200-bit offsets,
     then by random 64-bit offsets                                 (--advance)

With --clock (no profiler needed) it instead prints, per launch, the kernel
time (HIP events) and the shader clock sampled right after it, for a
sequence started after 2 s idle: six 2^28 float64 fills, then twelve rounds
of { fill, write probe with the fill's store pattern, hipMemsetAsync, 2^28
uint32 fill with the vector kernel } over the same 2 GiB -- which kernels
slow down while the power manager holds the clock down.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--advance", action="store_true")
    ap.add_argument("--clock", action="store_true")
    args = ap.parse_args()
    if args.clock:
        clock_interleave()
        return

    import mrg8_amd
    from mrg8_amd import hip, workloads

    n = 1 << 28
    src = hip.DeviceBuffer(1 << 30)
    dst = hip.DeviceBuffer(1 << 30)
    src.fill_byte(1)
    for _ in range(args.warmup + args.reps):
        hip.check(hip._load().hipMemcpy(hip.ctypes.c_void_p(dst.ptr), hip.ctypes.c_void_p(src.ptr), 1 << 30, 3),
                  "hipMemcpy D2D")
    hip.synchronize()
    src.free()
    dst.free()
    buf = hip.DeviceBuffer(n * 8)
    g = mrg8_amd.mrg8_cuda(workloads.SEED)
    for fmt, shape in ((mrg8_amd.F64, {}), (mrg8_amd.U32, {}), (mrg8_amd.U32, {"per_lane": 4}),
                       (mrg8_amd.F64, {"per_lane": 16, "block_threads": 1024, "tile": 2048})):
        for _ in range(args.warmup + args.reps):
            g.generate(buf, n, fmt, **shape)
        g.synchronize()
    buf.free()
    if args.batched:
        states, lengths = workloads.batched()
        total = int(lengths.sum())
        out = hip.DeviceBuffer(total * 8)
        starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
        b = mrg8_amd.Batch(max_streams=len(lengths))
        b.set([(states[i], 0, int(lengths[i]), out.ptr + 8 * int(starts[i]), mrg8_amd.F64)
               for i in range(len(lengths))])
        # the batch default (one tile per wave), then the persistent grid that
        # takes the tiles in address order
        for shape in ((1024, 0, 2048), (512, -2, 2048)):
            b.set_config(*shape)
            for _ in range(args.warmup + args.reps):
                b.run()
            hip.synchronize()
        out.free()
        steps = workloads.jump_steps()
        m = len(steps)
        out = hip.DeviceBuffer(m * 1024)
        s0 = mrg8_amd.seed_state(workloads.SEED)
        j = mrg8_amd.Batch(max_streams=m, max_tiles=m)
        j.set_config(256, 8, 256)
        j.set([(s0, int(steps[i]), 256, out.ptr + 1024 * i, mrg8_amd.U32) for i in range(m)])
        for _ in range(args.warmup + args.reps):
            j.run()
        hip.synchronize()
        out.free()
    if args.advance:
        rng = np.random.default_rng(7)
        ns = 1 << 20
        states = rng.integers(1, 2**31 - 1, size=(ns, 8), dtype=np.int64).astype(np.uint32)
        d_in, d_out = hip.DeviceBuffer(ns * 32), hip.DeviceBuffer(ns * 32)
        d_in.copy_from_host(states)
        for words in (4, 1):
            offs = rng.integers(0, 2**63, size=(ns, words), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
            if words == 4:
                offs[:, 3] &= np.uint64(0xFF)  # 200-bit offsets
            d_off = hip.DeviceBuffer(offs.nbytes)
            d_off.copy_from_host(offs)
            for _ in range(args.warmup + args.reps):
                mrg8_amd.advance_states(d_in, d_off, words, d_out, ns)
            hip.synchronize()
            d_off.free()
        d_in.free()
        d_out.free()
    print("profile workload done", flush=True)


def clock_interleave():
    import time

    import mrg8_amd
    from mrg8_amd import hip, probe

    mrg8_amd.capi.call("mrg8_init_device", 0)
    n = 1 << 28
    s = hip.Stream()
    buf = hip.DeviceBuffer(n * 8)
    g = mrg8_amd.mrg8_cuda(13579)
    seq = [("fill f64", lambda: g.generate(buf, n, mrg8_amd.F64, stream=s))] * 6
    seq += [("fill f64", lambda: g.generate(buf, n, mrg8_amd.F64, stream=s)),
            ("write probe", lambda: probe.write(buf.ptr, n, 1, 0, s)),
            ("memset", lambda: hip.memset_async(buf.ptr, 0, n * 8, s)),
            ("fill u32", lambda: g.generate(buf, n, mrg8_amd.U32, stream=s, per_lane=4))] * 12
    clk = probe.ClockProbe(len(seq))
    ev = [(hip.Event(), hip.Event()) for _ in seq]
    s.synchronize()
    time.sleep(2.0)
    for (name, fn), (e0, e1) in zip(seq, ev):
        e0.record(s)
        fn()
        e1.record(s)
        clk.sample(s)
    s.synchronize()
    mhz = clk.mhz()
    print("launch  kernel        us    MHz")
    for i, ((name, _), (e0, e1)) in enumerate(zip(seq, ev)):
        print(f"{i:6d}  {name:11s} {e0.elapsed_ms(e1) * 1e3:7.1f} {mhz[i]:6.0f}", flush=True)
    clk.free()
    buf.free()


if __name__ == "__main__":
    main()
