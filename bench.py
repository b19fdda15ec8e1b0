"""MRG8 benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]

A step is one generation of 2^28 float64 outputs (x * (1/(2^31-1)), the
reference's `rand` conversion) of a single stream seeded with 13579, split
across all waves by jump-ahead -- the single-stream workload of the
benchmark configurations.  Printed value: G samples/s over the timed steps.

The same JSON line carries, measured in the same run (in this order):
  write_ceiling the project's write probe (d[i] = i, the counterpart of the
                reference's cuda/stream.cu) over the same 2 GiB with the fill's
                store pattern (and its runs in dispatch order) and in the
                reference's form, and hipMemsetAsync
  roofline      algorithmic bytes per launch / mean kernel time (HIP events on
                the launch stream, one between every two launches; every
                launch's duration listed) against 8 TB/s; HBM traffic from the
                committed rocprofv3 counter summary when present
  u32           the same workload in raw uint32 words (the matrix-core fill),
                and with the vector kernel
  f64_matrix_core  the float64 workload with the matrix-core fill
  small         2^24 float64 outputs (the reference's full-compare size)
  clock         the float64 fill from idle with the shader clock sampled after
                each launch (the power manager's clamp), then the uint32 fill
  batched       4096 streams, log-uniform lengths in [2^10, 2^22), float64, one
                launch with the autotuned shape; also as one replay of a
                captured HIP graph (persistent grids taking the tiles in
                address order or striding over them, one tile per wave)
  jumps         65536 streams jumped by uniform 63-bit steps, 256 uint32 each,
                with the multiply-add rate against the measured VALU peak
  advance       2^20 device-resident states jumped by 200-bit (and 64-bit)
                offsets (mrg8_advance_states)
  served        1e5 Poisson requests on 16384 resident streams, sizes
                log-uniform in [2^8, 2^26) float64, 64 GB admission cap:
                throughput and p50/p99 latency (arrival to completion)
  cpu_baseline  the reference's serial generator (original/mrg8.cpp, compiled
                from its sources) on one host core, bounded sample
  cpu_baseline_vec  the knl vector/OpenMP restatement on all host cores
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G samples/s & % HBM write roofline (1 stream, batched); p99 latency served"
PEAK_GBS = 8000.0
N_SINGLE = 1 << 28
SEED = 13579


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_launches(fn, stream, reps):
    """Mean ms per launch: HIP events on `stream` bracketing `reps` back-to-back launches."""
    from mrg8_amd import hip

    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    e0.destroy()
    e1.destroy()
    return ms


def traffic_from_profiles(name):
    """HBM bytes per launch recorded by the committed counter summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def valu_busy_from_profiles(name, kernel):
    """VALUBusy (%) of `kernel` in the committed counter summary, if recorded."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 3 and parts[0] == kernel and parts[1] == "VALUBusy":
            try:
                return float(parts[2])
            except ValueError:
                return None
    return None


def timed_series(step, stream, warmup, steps):
    """W untimed launches, then K launches with a HIP event between every two
    (on the launch stream): wall seconds, mean ms per launch and the K
    per-launch durations."""
    from mrg8_amd import hip

    for _ in range(warmup):
        step()
    stream.synchronize()
    hip.synchronize()
    ev = [hip.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        step()
        ev[i + 1].record(stream)
    ev[-1].synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t0
    per = [ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)]
    total = ev[0].elapsed_ms(ev[-1])
    for e in ev:
        e.destroy()
    return wall, total / steps, per


def write_ceiling(buf, nbytes, s, reps=10):
    """Store-bandwidth ceiling over the same buffer: this project's probe
    (d[i] = i, the counterpart of cuda/stream.cu) with the double fill's 1 KiB
    non-temporal wave stores (512-thread blocks in XCD runs), with the same
    runs in 1024-thread blocks in dispatch order, and in the reference's
    one-element-per-thread form, and the runtime's own fill kernel
    (hipMemsetAsync)."""
    from mrg8_amd import hip, probe

    n = nbytes // 8
    out = {}
    for name, fn in (("probe_fill_pattern", lambda: probe.write(buf.ptr, n, 1, 0, s)),
                     ("probe_dispatch_order", lambda: probe.write(buf.ptr, n, 2, 0, s)),
                     ("probe_reference_form", lambda: probe.write(buf.ptr, n, 0, 0, s)),
                     ("memset", lambda: hip.memset_async(buf.ptr, 0, nbytes, s))):
        fn()
        km = time_launches(fn, s, reps)
        out[name] = {"ms": round(km, 4), "gbs": round(nbytes / (km * 1e-3) / 1e9, 1)}
    log("write ceiling: " + ", ".join(f"{k} {v['gbs']:.0f} GB/s" for k, v in out.items()))
    return out


def bench_single(args):
    import mrg8_amd
    from mrg8_amd import hip

    s = hip.Stream()
    buf = hip.DeviceBuffer(N_SINGLE * 8)
    g = mrg8_amd.mrg8_cuda(SEED)
    out = {"ceiling": write_ceiling(buf, N_SINGLE * 8, s)}
    # default shapes (doubles: the vector kernel; words: the matrix-core
    # fill), then each format with the other kernel for comparison
    legs = (("f64", mrg8_amd.F64, 8, {}), ("u32", mrg8_amd.U32, 4, {}),
            ("u32_vector", mrg8_amd.U32, 4, {"per_lane": 4}),
            ("f64_matrix_core", mrg8_amd.F64, 8, {"per_lane": 16, "block_threads": 1024, "tile": 2048}))
    for name, fmt, esz, shape in legs:
        step = lambda: g.generate(buf, N_SINGLE, fmt, stream=s, **shape)  # noqa: E731
        wall, km, per = timed_series(step, s, args.warmup, args.steps)
        out[name] = {"wall_s": wall, "kernel_ms": km, "per_launch_ms": per, "bytes": N_SINGLE * esz}
        log(f"single {name}: {N_SINGLE * args.steps / wall / 1e9:.1f} G samples/s wall, kernel {km:.3f} ms "
            f"(launches {min(per):.3f}-{max(per):.3f} ms)")
    # the reference's full-compare size
    small = 1 << 24
    step = lambda: g.generate(buf, small, mrg8_amd.F64, stream=s)  # noqa: E731
    for _ in range(3):
        step()
    km = time_launches(step, s, 20)
    out["small"] = {"kernel_ms": km, "gsps": small / (km * 1e-3) / 1e9,
                    "gbs": small * 8 / (km * 1e-3) / 1e9}
    out["clock"] = clock_series(g, buf, s)
    # the word fill again right after that series, with a clock sample after
    # each launch: the same kernel once the clock has recovered (its timed
    # window above follows the float64 leg, inside the clamp)
    out["u32_after_clock_series"] = clock_tail(g, buf, s, mrg8_amd.U32, 30)
    buf.free()
    s.destroy()
    return out


def clock_tail(g, buf, s, fmt, launches):
    from mrg8_amd import hip, probe

    clk = probe.ClockProbe(launches)
    ev = [hip.Event() for _ in range(2 * launches)]
    for i in range(launches):
        ev[2 * i].record(s)
        g.generate(buf, N_SINGLE, fmt, stream=s)
        ev[2 * i + 1].record(s)
        clk.sample(s)
    s.synchronize()
    us = [ev[2 * i].elapsed_ms(ev[2 * i + 1]) * 1e3 for i in range(launches)]
    mhz = clk.mhz()
    clk.free()
    for e in ev:
        e.destroy()
    k = us[-15:]
    return {"launches": launches, "last15_kernel_ms": round(sum(k) / len(k) * 1e-3, 4),
            "last15_sclk_mhz": round(sum(mhz[-15:]) / 15, 1), "min_sclk_mhz": min(mhz)}


def clock_series(g, buf, s, idle_s=2.0, launches=60):
    """The 2^28-double fill from an idle GPU, each launch followed by a
    shader-clock sample: how the power manager's clock follows a sustained
    fill (the timed window above starts the same way)."""
    import mrg8_amd
    from mrg8_amd import hip, probe

    clk = probe.ClockProbe(launches)
    ev = [hip.Event() for _ in range(2 * launches)]
    s.synchronize()
    time.sleep(idle_s)
    for i in range(launches):
        ev[2 * i].record(s)
        g.generate(buf, N_SINGLE, mrg8_amd.F64, stream=s)
        ev[2 * i + 1].record(s)
        clk.sample(s)
    s.synchronize()
    us = [round(ev[2 * i].elapsed_ms(ev[2 * i + 1]) * 1e3, 1) for i in range(launches)]
    mhz = clk.mhz()
    clk.free()
    for e in ev:
        e.destroy()
    tail = us[-15:]
    steady = sum(tail) / len(tail) * 1e-3
    log(f"clock series: {min(mhz):.0f}-{max(mhz):.0f} MHz, fill {min(us):.0f}-{max(us):.0f} us, "
        f"last 15 launches {steady:.3f} ms")
    return {"idle_before_s": idle_s, "kernel_us": us, "sclk_mhz": mhz,
            "last15_kernel_ms": round(steady, 4),
            "last15_frac": round(N_SINGLE * 8 / (steady * 1e-3) / 1e9 / PEAK_GBS, 4),
            "last15_sclk_mhz": round(sum(mhz[-15:]) / 15, 1)}


def bench_host_form(n=1 << 26, reps=3):
    """mrg8_rand into pageable host memory (device generation + D2H copy):
    the PCIe-inclusive rate a host-buffer caller sees; never the headline."""
    import ctypes

    import mrg8_amd
    from mrg8_amd import capi

    g = mrg8_amd.mrg8_cuda(SEED)
    out = np.empty(n, dtype=np.float64)
    out.fill(0.0)  # touch the pages before timing
    times = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        capi.call("mrg8_rand", g._h, ctypes.c_void_p(out.ctypes.data), ctypes.c_uint64(n))
        times.append(time.perf_counter() - t0)
    g.close()
    mean = sum(times[1:]) / reps
    log(f"host form (mrg8_rand, pageable, PCIe-inclusive): {n / mean / 1e9:.2f} G samples/s")
    return {"outputs": n, "ms": mean * 1e3, "gsps": n / mean / 1e9, "gbs": n * 8 / mean / 1e9,
            "destination": "pageable host memory (numpy)"}


def bench_batched(args):
    import mrg8_amd
    from mrg8_amd import hip, workloads

    states, lengths = workloads.batched()
    total = int(lengths.sum())
    buf = hip.DeviceBuffer(total * 8)
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    b = mrg8_amd.Batch(max_streams=len(lengths))
    b.set([(states[i], 0, int(lengths[i]), buf.ptr + 8 * int(starts[i]), mrg8_amd.F64) for i in range(len(lengths))])
    s = hip.Stream()
    cfg = b.autotune(stream=s, cache_path=os.path.join(tempfile.gettempdir(), "mrg8_batch_tune.txt"))
    for _ in range(2):
        b.run(s)
    km = time_launches(lambda: b.run(s), s, max(5, args.steps))
    gbs = total * 8 / (km * 1e-3) / 1e9
    log(f"batched: {total / (km * 1e-3) / 1e9:.1f} G samples/s, {gbs:.0f} GB/s, cfg {cfg}")
    out = {"streams": len(lengths), "outputs": total, "ms_per_launch": km, "gsps": total / (km * 1e-3) / 1e9,
           "gbs": gbs, "frac_of_peak": gbs / PEAK_GBS, "config": cfg,
           "launch": "direct launch, " + ("one tile per wave" if cfg["blocks_per_cu"] == 0 else
                                          "persistent grid taking the tiles in address order"
                                          if cfg["blocks_per_cu"] < 0 else "persistent grid striding over the tiles")}
    # the configuration's named form: one replay of a captured HIP graph per
    # run, with a persistent grid taking the tiles in address order, with one
    # striding over them (best of a few shapes each) and with one tile per wave
    b.use_graph(True)
    for label, shapes in (("queue_graph", [(512, -2, 2048), (512, -3, 2048), (1024, -1, 2048), (256, -4, 2048)]),
                          ("persistent_graph", [(256, 4, 4096), (256, 8, 4096), (512, 2, 4096), (256, 4, 8192),
                                                (1024, 1, 4096), (256, 8, 2048)]),
                          ("tiles_graph", [(1024, 0, 2048)])):
        best = None
        for shape in shapes:
            b.set_config(*shape)
            for _ in range(2):
                b.run(s)
            t = time_launches(lambda: b.run(s), s, max(5, args.steps))
            if best is None or t < best[1]:
                best = (shape, t)
        g = total * 8 / (best[1] * 1e-3) / 1e9
        out[label] = {"ms_per_launch": round(best[1], 4), "gsps": round(total / (best[1] * 1e-3) / 1e9, 2),
                      "gbs": round(g, 1), "frac_of_peak": round(g / PEAK_GBS, 4),
                      "config": dict(zip(("block_threads", "blocks_per_cu", "tile_outputs"), best[0])),
                      "shapes_tried": len(shapes)}
        log(f"batched {label}: {best[1]:.3f} ms, {g:.0f} GB/s, shape {best[0]}")
    b.use_graph(False)
    buf.free()
    return out


def bench_jumps(args):
    import mrg8_amd
    from mrg8_amd import hip, workloads

    steps = workloads.jump_steps()
    n = len(steps)
    buf = hip.DeviceBuffer(n * 256 * 4)
    s0 = mrg8_amd.seed_state(SEED)
    b = mrg8_amd.Batch(max_streams=n, max_tiles=n)
    b.set_config(256, 8, 256)
    b.set([(s0, int(steps[i]), 256, buf.ptr + 1024 * i, mrg8_amd.U32) for i in range(n)])
    s = hip.Stream()
    for _ in range(2):
        b.run(s)
    km = time_launches(lambda: b.run(s), s, max(5, args.steps))
    buf.free()
    pc = float(np.mean([bin(int(x)).count("1") for x in steps]))
    digits = float(np.mean([sum(1 for k in range(8) if (int(x) >> (8 * k)) & 0xFF) for x in steps]))
    # multiply-adds: the reference's jump (one 8x8 matrix-vector product, 64
    # MACs, per set bit, knl/mrg8.h:212-238) plus 8 per output; and what the
    # kernel executes (one product per non-zero byte of the step)
    macs_ref = n * (pc * 64 + 256 * 8)
    macs_run = n * (digits * 64 + 256 * 8)
    hbm = n * (256 * 4 + 64)  # outputs + one 64-byte segment record per stream
    log(f"jumps: {km:.3f} ms per launch, {n / (km * 1e-3) / 1e6:.1f} M streams/s, "
        f"{macs_run / (km * 1e-3) / 1e12:.2f} T MACs/s")
    return {"streams": n, "outputs_per_stream": 256, "ms_per_launch": km, "streams_per_s": n / (km * 1e-3),
            "gsps": n * 256 / (km * 1e-3) / 1e9, "mean_popcount": pc, "mean_nonzero_bytes": digits,
            "jump_macs_per_s": n * pc * 64 / (km * 1e-3),
            "valu": {"bound": "valu", "unit": "lane MACs/s (32x32->64)",
                     "executed_macs_per_launch": macs_run, "reference_algorithm_macs_per_launch": macs_ref,
                     "achieved": round(macs_run / (km * 1e-3), 1)},
            "hbm": {"bytes_per_launch": hbm, "gbs": round(hbm / (km * 1e-3) / 1e9, 1),
                    "frac": round(hbm / (km * 1e-3) / 1e9 / PEAK_GBS, 4)}}


def bench_advance(args, n=1 << 20):
    """mrg8_advance_states: n device-resident states, each jumped by its own
    random 200-bit offset (the reference's widest jump, done one state at a
    time on the host there), and by 64-bit offsets for comparison."""
    import mrg8_amd
    from mrg8_amd import hip

    rng = np.random.default_rng(7)
    states = rng.integers(1, 2**31 - 1, size=(n, 8), dtype=np.int64).astype(np.uint32)
    d_in, d_out = hip.DeviceBuffer(n * 32), hip.DeviceBuffer(n * 32)
    d_in.copy_from_host(states)
    out = {"states": n}
    s = hip.Stream()
    for words, bits in ((4, 200), (1, 64)):
        offs = rng.integers(0, 2**63, size=(n, words), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        if bits == 200:
            offs[:, 3] &= np.uint64(0xFF)  # bits 200..255 clear
        d_off = hip.DeviceBuffer(offs.nbytes)
        d_off.copy_from_host(offs)
        for _ in range(3):
            mrg8_amd.advance_states(d_in, d_off, words, d_out, n, stream=s)
        km = time_launches(lambda: mrg8_amd.advance_states(d_in, d_off, words, d_out, n, stream=s), s, 20)
        d_off.free()
        out[f"{bits}bit_ms"] = km
        out[f"{bits}bit_states_per_s"] = n / (km * 1e-3)
        log(f"advance: {n} states by {bits}-bit offsets in {km:.3f} ms ({n / (km * 1e-3) / 1e9:.2f} G states/s)")
    d_in.free()
    d_out.free()
    return out


def bench_served(args, capacity_gsps):
    import mrg8_amd
    from mrg8_amd import workloads

    n_req = args.served_requests
    mean = workloads.mean_served_size()
    rate = args.served_load * capacity_gsps * 1e9 / mean  # requests per second at the offered load
    states, arrivals, ids, sizes = workloads.served(n_requests=n_req, rate_per_s=rate)
    srv = mrg8_amd.Server(num_streams=len(states), max_output_bytes=64 << 30, max_inflight=4096,
                          step_outputs=args.step_outputs)
    srv.set_streams(states)
    # warm-up trace (separate streams' positions advance; states stay resident)
    srv.replay(arrivals[:200] / 10, ids[:200], sizes[:200], fmt=mrg8_amd.F64)
    c0 = srv.counters()
    stats, lat = srv.replay(arrivals, ids, sizes, fmt=mrg8_amd.F64, want_latencies=True)
    c1 = srv.counters()
    srv.close()
    # where in the trace the slow requests arrived: p99 per 100 ms of arrivals
    windows = []
    for w0 in np.arange(0.0, float(arrivals[-1]) + 1e-9, 0.1):
        m = (arrivals >= w0) & (arrivals < w0 + 0.1)
        if m.any():
            windows.append(round(float(np.quantile(lat[m], 0.99)) * 1e3, 3))
    worst = int(np.argmax(lat))
    steps = max(1, c1["steps"] - c0["steps"])
    out = {"requests": n_req, "streams": len(states), "offered_load": args.served_load,
           "arrival_rate_per_s": rate, "outputs": int(stats["outputs"]), "wall_s": stats["wall_s"],
           "gsps": stats["outputs"] / stats["wall_s"] / 1e9, "p50_ms": stats["p50_s"] * 1e3,
           "p90_ms": stats["p90_s"] * 1e3, "p99_ms": stats["p99_s"] * 1e3, "max_ms": stats["max_s"] * 1e3,
           "admission_cap_bytes": 64 << 30,
           "steps": steps, "mean_step_outputs": (c1["step_outputs"] - c0["step_outputs"]) / steps,
           "mean_step_segments": (c1["step_segments"] - c0["step_segments"]) / steps,
           "host_us_per_step": (c1["host_launch_s"] - c0["host_launch_s"]) / steps * 1e6,
           "max_admission_wait_ms": c1["max_queue_s"] * 1e3,
           "p99_ms_per_100ms_of_arrivals": windows,
           "worst": {"arrival_ms": round(float(arrivals[worst]) * 1e3, 3), "latency_ms": round(float(lat[worst]) * 1e3, 3),
                     "outputs": int(sizes[worst])},
           "max_host_gap_ms": round(stats["max_host_gap_s"] * 1e3, 3),
           "max_host_gap_at_ms": round(stats["max_host_gap_at_s"] * 1e3, 3)}
    log(f"served: {out['gsps']:.1f} G samples/s, p50 {out['p50_ms']:.3f} ms, p99 {out['p99_ms']:.3f} ms")
    return out


def cpu_baselines(args):
    """The oracle/_ref reference build (1 core) and the vector port (all cores)."""
    import oracle

    res = {}
    n = 1 << args.cpu_log2
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_original")
    if os.path.exists(ref) and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "jump_matrix.txt")):
        r = subprocess.run([ref], cwd=os.path.dirname(ref), input=f"T {SEED} {n} 10\nQ\n", capture_output=True,
                           text=True, timeout=600)
        ms = [float(x) for x in r.stdout.split()]
        mean = float(np.mean(ms[1:]))  # the reference's method: drop the first of 10, mean of 9
        res["cpu_baseline"] = {"value": n / (mean * 1e-3) / 1e9, "unit": "G samples/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"original/mrg8.cpp rand(double*, 2^{args.cpu_log2}) from seed {SEED}, "
                                         f"10 runs, first dropped, mean of 9 = {mean:.1f} ms"}
    else:
        st = oracle.seed_state(SEED)
        buf = np.empty(n, dtype=np.float64)
        ts = []
        for _ in range(4):
            t0 = time.perf_counter()
            oracle.lib().mrg8o_rand(oracle._ptr(st.copy()), buf.ctypes.data, n)
            ts.append(time.perf_counter() - t0)
        mean = float(np.mean(ts[1:]))
        res["cpu_baseline"] = {"value": n / mean / 1e9, "unit": "G samples/s", "cores": 1, "kind": "port",
                               "sample": f"oracle serial rand, 2^{args.cpu_log2} outputs, mean of 3"}
    # the reference's 200-bit jump (J = product of table matrices, then J x state), one state at a time
    rng = np.random.default_rng(7)
    nj = 2000
    offs = [int.from_bytes(rng.bytes(25), "little") for _ in range(nj)]
    st = oracle.seed_state(SEED)
    t0 = time.perf_counter()
    for x in offs:
        oracle.jump_bits(st, [b for b in range(200) if (x >> b) & 1])
    dt = time.perf_counter() - t0
    res["cpu_baseline_advance"] = {"value": nj / dt, "unit": "states/s", "cores": 1, "kind": "port",
                                   "sample": f"oracle restatement of original/mrg8.cpp:105-122 (jump_ahead(short[200])), "
                                             f"{nj} random 200-bit offsets, one pass"}
    # vector + OpenMP restatement of the knl path on every host core
    nv = 1 << 28
    buf = np.empty(nv, dtype=np.float64)
    st = oracle.seed_state(SEED)
    ts, used = [], 0
    for _ in range(10):
        t0 = time.perf_counter()
        _, _, used = oracle.vec_outer_tp(st, nv, 0, out=buf)
        ts.append(time.perf_counter() - t0)
    mean = float(np.mean(ts[1:]))
    res["cpu_baseline_vec"] = {"value": nv / mean / 1e9, "unit": "G samples/s", "cores": used, "kind": "port",
                               "avx512": oracle.vec_has_avx512(),
                               "sample": "knl/mrg8_vec.h outer-product restatement, 2^28 (x-1)/p doubles, "
                                         "10 runs, first dropped, mean of 9"}
    # the 4096-stream batch, OpenMP over streams (each stream generated serially
    # in cache-sized blocks and checksummed, not stored)
    from mrg8_amd import workloads

    states, lengths = workloads.batched()
    zeros = np.zeros(len(lengths), dtype=np.uint64)
    fmts = np.zeros(len(lengths), dtype=np.int32)
    t0 = time.perf_counter()
    oracle.stream_wsums(states, zeros, lengths, fmts)
    wall = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    res["cpu_baseline_batched"] = {"value": float(lengths.sum()) / wall / 1e9, "unit": "G samples/s",
                                   "cores": threads, "kind": "port",
                                   "sample": "all 4096 streams of the batched workload (2.13e9 float64), "
                                             "oracle serial recurrence per stream, OpenMP over streams, "
                                             "outputs checksummed in 4096-value blocks instead of stored; one run"}
    log(f"cpu batched: {res['cpu_baseline_batched']['value']:.2f} G samples/s on {threads} threads")
    # the jump-heavy batch: 65536 x (63-bit jump by the knl matrix-vector chain
    # + 256 words), OpenMP over streams
    steps = workloads.jump_steps()
    m = len(steps)
    st = np.tile(oracle.seed_state(SEED), (m, 1))
    t0 = time.perf_counter()
    oracle.stream_wsums(st, steps, np.full(m, 256, dtype=np.uint64), np.full(m, 2, dtype=np.int32))
    wall = time.perf_counter() - t0
    res["cpu_baseline_jumps"] = {"value": wall * 1e3, "unit": "ms", "cores": threads, "kind": "port",
                                 "sample": "the whole jump-heavy batch (65536 streams: jump by the matrix-vector chain "
                                           "of knl/mrg8.h:212-238, then 256 words each, checksummed), OpenMP over "
                                           "streams; one run"}
    log(f"cpu jumps: {wall * 1e3:.1f} ms on {threads} threads")
    # the served mix: the first 5000 requests (about 2.7e10 float64), each a
    # jump to its reserved position + its outputs, OpenMP over requests
    states, _, ids, sizes = workloads.served()
    k = 5000
    cursor = np.zeros(len(states), dtype=np.uint64)
    offs = np.empty(k, dtype=np.uint64)
    for i in range(k):
        offs[i] = cursor[ids[i]]
        cursor[ids[i]] += sizes[i]
    t0 = time.perf_counter()
    oracle.stream_wsums(states[ids[:k]], offs, sizes[:k], np.zeros(k, dtype=np.int32))
    wall = time.perf_counter() - t0
    res["cpu_baseline_served"] = {"value": float(sizes[:k].sum()) / wall / 1e9, "unit": "G samples/s",
                                  "cores": threads, "kind": "port",
                                  "sample": f"first {k} requests of the served mix ({float(sizes[:k].sum()):.3g} "
                                            "float64), throughput only (no arrivals), OpenMP over requests; one run"}
    log(f"cpu served: {res['cpu_baseline_served']['value']:.2f} G samples/s on {threads} threads")
    try:
        res["host_cpu"] = subprocess.run(["sh", "-c", "grep -m1 'model name' /proc/cpuinfo | cut -d: -f2"],
                                         capture_output=True, text=True).stdout.strip()
    except Exception:
        pass
    res["host_nproc"] = os.cpu_count()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--quick", action="store_true", help="single-stream figures only")
    ap.add_argument("--served-requests", type=int, default=100_000)
    ap.add_argument("--served-load", type=float, default=0.7)
    ap.add_argument("--step-outputs", type=int, default=0, help="server step size (0 = library default)")
    ap.add_argument("--cpu-log2", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-comparators", action="store_true", help="skip the rocRAND context timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the generator needs no "
              f"cross-GPU exchange: each GPU would fill its own range of the sequence, its start state "
              f"jumped ahead on the host)", file=sys.stderr)
        sys.exit(2)

    import mrg8_amd
    from mrg8_amd import hip

    if hip.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    mrg8_amd.capi.call("mrg8_init_device", 0)

    single = bench_single(args)
    f64 = single["f64"]
    value = N_SINGLE * args.steps / f64["wall_s"] / 1e9
    achieved = f64["bytes"] / (f64["kernel_ms"] * 1e-3) / 1e9
    traffic = traffic_from_profiles("fill_f64_2e28_counters.txt")
    ceiling = single["ceiling"]["probe_fill_pattern"]
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "G samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f64["wall_s"] / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "single stream, seed 13579, 2^28 float64 outputs double(x)*(1/(2^31-1)), split "
                               "across all waves by jump-ahead (exact 64-bit integer arithmetic mod 2^31-1)",
                   "outputs": N_SINGLE, "output_dtype": "float64", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_fill<F64,4>",
                     "shape": "default: 512-thread blocks, 2048 outputs per wave, 4 per lane, blocks in XCD runs "
                              "of 64 (each XCD writes 64 consecutive blocks' output)",
                     "kernel_ms": round(f64["kernel_ms"], 4),
                     "per_launch_ms": [round(x, 4) for x in f64["per_launch_ms"]],
                     "valu_busy_pct": valu_busy_from_profiles("fill_f64_2e28_counters.txt", "k_fill<F64,4>"),
                     "algorithmic_bytes_per_launch": f64["bytes"]},
        "u32": {"gsps": round(N_SINGLE * args.steps / single["u32"]["wall_s"] / 1e9, 2),
                "kernel": "k_fill_mfma<U32,128,8>: the matrix-core fill (int8 MFMA), the default for words; "
                          "512-thread blocks, 8192 words per wave",
                "kernel_ms": round(single["u32"]["kernel_ms"], 4),
                "vector_kernel_ms": round(single["u32_vector"]["kernel_ms"], 4),
                "vector_per_launch_ms": [round(x, 4) for x in single["u32_vector"]["per_launch_ms"]],
                "gbs": round(single["u32"]["bytes"] / (single["u32"]["kernel_ms"] * 1e-3) / 1e9, 1),
                "frac": round(single["u32"]["bytes"] / (single["u32"]["kernel_ms"] * 1e-3) / 1e9 / PEAK_GBS, 4),
                "per_launch_ms": [round(x, 4) for x in single["u32"]["per_launch_ms"]],
                "after_clock_series": dict(single["u32_after_clock_series"], frac=round(
                    single["u32"]["bytes"] / single["u32_after_clock_series"]["last15_kernel_ms"] / 1e6 / PEAK_GBS,
                    4))},
        "f64_matrix_core": {
            "kernel": "k_fill_mfma<F64,64,16>: the matrix-core fill for doubles (outputs_per_lane 16, "
                      "1024-thread blocks, 2048 per wave), measured after the word legs",
            "kernel_ms": round(single["f64_matrix_core"]["kernel_ms"], 4),
            "frac": round(single["f64_matrix_core"]["bytes"] / (single["f64_matrix_core"]["kernel_ms"] * 1e-3) / 1e9
                          / PEAK_GBS, 4),
            "per_launch_ms": [round(x, 4) for x in single["f64_matrix_core"]["per_launch_ms"]]},
        "small_2e24_f64": {k: round(v, 4) for k, v in single["small"].items()},
        "write_ceiling": {"kernel": "mrg8_probe_write: d[i] = i over the same 2 GiB with the fill's 1 KiB "
                                    "non-temporal wave stores, 512-thread blocks in XCD runs of 64 (the "
                                    "counterpart of cuda/stream.cu), measured just before the float64 leg",
                          "gbs": ceiling["gbs"], "frac": round(ceiling["gbs"] / PEAK_GBS, 4),
                          "fill_frac_of_ceiling": round(achieved / ceiling["gbs"], 4),
                          "dispatch_order_gbs": single["ceiling"]["probe_dispatch_order"]["gbs"],
                          "reference_form_gbs": single["ceiling"]["probe_reference_form"]["gbs"],
                          "memset_gbs": single["ceiling"]["memset"]["gbs"]},
        "clock": dict(single["clock"], note=(
            "shader clock (s_memtime against the 100 MHz s_memrealtime, sampled after each launch) of a "
            "separate 60-launch series of the same fill started after 2 s idle: the power manager lowers "
            "the clock a few launches into a sustained fill and restores it over the next tens of launches; "
            "the timed window above (launches warmup+1..warmup+steps) falls in that span when warmup is small, "
            "and last15_* is the same fill once the clock has recovered")),
    }
    if not args.quick:
        line["batched"] = bench_batched(args)
        line["jumps"] = bench_jumps(args)
        line["advance"] = bench_advance(args)
        line["served"] = bench_served(args, line["batched"]["gsps"])
        line["host_form_f64"] = bench_host_form()
    if not args.quick and not args.no_comparators:
        from mrg8_amd import comparators

        rates = comparators.rocrand_rates(N_SINGLE)
        line["rocrand_f64_2e28"] = {k: {kk: round(vv, 4) for kk, vv in v.items()} if "error" not in v else v
                                    for k, v in rates.items()}
        log("rocrand: " + ", ".join(f"{k} {v.get('steady_gsps', 0):.1f}" for k, v in rates.items()))
    if not args.quick:
        # the VALU ceiling of the jump-heavy batch, measured last: a
        # multiply-add-only burst lowers the clock for the launches after it
        from mrg8_amd import hip, probe

        s = hip.Stream()
        peak = probe.mad64_rate(s)
        s.destroy()
        v = line["jumps"]["valu"]
        v.update(peak_measured=round(peak, 1), frac=round(v["achieved"] / peak, 4),
                 peak_kernel="mrg8_probe_mad64: 8 independent v_mad_i64_i32 chains per lane, run after every GPU leg")
        log(f"jumps VALU: {v['achieved'] / 1e12:.2f} of {peak / 1e12:.1f} T MACs/s measured peak")
    if not args.no_cpu:
        line.update(cpu_baselines(args))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
