"""The served configuration at its full size: 16384 device-resident streams,
1e5 Poisson requests of log-uniform 2^8..2^26 float64 outputs, a 64 GB
admission cap.  Every request's first and last 64 outputs are compared with
the oracle at the request's reserved stream position, a fixed-seed 1 % sample
of the requests of at most 2^20 outputs is compared in full, the cap is
checked after every poll, and three sampled requests are served again alone
by a fresh server and must give the same bytes.

The arrival rate is lower than the benchmark's (a Python loop copies the
edges of every request back before releasing it); the workload is otherwise
the benchmark's (mrg8_amd.workloads.served)."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = 64  # outputs checked at each end of every request


def d2h(dst: np.ndarray, src_addr: int, nbytes: int) -> None:
    from mrg8_amd import hip

    hip.check(hip._load().hipMemcpy(ctypes.c_void_p(dst.ctypes.data), ctypes.c_void_p(src_addr), nbytes, 2), "D2H")


def expected_starts(ids, sizes, n_streams):
    """Stream positions reserved at submit, in arrival order."""
    starts = np.zeros(len(ids), dtype=np.uint64)
    cursor = np.zeros(n_streams, dtype=np.uint64)
    for i, (s, c) in enumerate(zip(ids.tolist(), sizes.tolist())):
        starts[i] = cursor[s]
        cursor[s] += np.uint64(c)
    return starts


def test_c5_served_full_workload(gpu, oracle_mod):
    from mrg8_amd import workloads

    n, n_streams, cap = 100_000, 16384, 64 << 30
    rate = 20_000.0
    states, arrivals, ids, sizes = workloads.served(n_requests=n, n_streams=n_streams, rate_per_s=rate)
    rng = np.random.default_rng(45)
    full_pick = (sizes <= (1 << 20)) & (rng.random(n) < 0.01)
    print(f"[case] c5 served: {n} requests, {n_streams} streams, {int(sizes.sum())} outputs, "
          f"{int(full_pick.sum())} compared in full", flush=True)
    srv = gpu.Server(num_streams=n_streams, max_output_bytes=cap, max_inflight=4096)
    srv.set_streams(states)
    heads = np.empty((n, K), dtype=np.float64)
    tails = np.empty((n, K), dtype=np.float64)
    starts = np.zeros(n, dtype=np.uint64)
    full = {}
    index = {}
    nxt = fin = 0
    max_bytes = 0
    t0 = time.perf_counter()
    while fin < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            index[srv.submit(int(ids[nxt]), int(sizes[nxt]), 0)] = nxt
            nxt += 1
        done = srv.progress(max_done=4096, wait_ms=0 if nxt < n else 2)
        st = srv.stats()  # completed requests not yet released still hold their bytes
        assert st["bytes_inflight"] <= cap
        max_bytes = max(max_bytes, st["bytes_inflight"])
        for c in done:
            i = index.pop(c["request_id"])
            cnt = int(c["count"])
            assert cnt == int(sizes[i]) and c["stream_id"] == int(ids[i]) and c["status"] == 0
            starts[i] = c["start"]
            d2h(heads[i], c["out"], K * 8)
            d2h(tails[i], c["out"] + (cnt - K) * 8, K * 8)
            if full_pick[i]:
                buf = np.empty(cnt, dtype=np.float64)
                d2h(buf, c["out"], cnt * 8)
                full[i] = buf
            srv.release(c["request_id"])
            fin += 1
            if fin % 25_000 == 0:
                print(f"[c5] {fin} requests checked out after {time.perf_counter() - t0:.1f} s", flush=True)
    wall = time.perf_counter() - t0
    st = srv.stats()
    assert st["completed"] == n and st["queued"] == 0 and st["bytes_inflight"] == 0
    print(f"[c5] served in {wall:.1f} s, most arena bytes held at once {max_bytes}", flush=True)

    # positions reserved at submit, in arrival order, per stream
    assert np.array_equal(starts, expected_starts(ids, sizes, n_streams))
    base = states[ids]
    want_h = oracle_mod.stream_windows(base, starts, K, 0)
    want_t = oracle_mod.stream_windows(base, starts + sizes - np.uint64(K), K, 0)
    bad_h = np.flatnonzero((heads.view(np.uint64) != want_h.view(np.uint64)).any(axis=1))
    bad_t = np.flatnonzero((tails.view(np.uint64) != want_t.view(np.uint64)).any(axis=1))
    assert bad_h.size == 0 and bad_t.size == 0, (bad_h[:5], bad_t[:5])
    for i, got in full.items():
        want, _ = oracle_mod.generate(oracle_mod.jump(base[i], int(starts[i])), int(sizes[i]), 0)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), i
    srv.close()

    # three of the sampled requests, served alone by a fresh server from the
    # same stream state and position, give the bytes they gave in the mix
    probes = sorted(full, key=lambda i: -int(sizes[i]))[:3]
    alone = gpu.Server(num_streams=n_streams, max_output_bytes=1 << 30)
    for i in probes:
        alone.set_stream(int(ids[i]), states[ids[i]], int(starts[i]))
        rid = alone.submit(int(ids[i]), int(sizes[i]), 0)
        got = []
        while not got:
            got = alone.progress(max_done=1, wait_ms=100)
        c = got[0]
        assert c["request_id"] == rid and int(c["start"]) == int(starts[i])
        buf = np.empty(int(sizes[i]), dtype=np.float64)
        d2h(buf, c["out"], buf.nbytes)
        alone.release(rid)
        assert np.array_equal(buf.view(np.uint64), full[i].view(np.uint64)), i
    alone.close()
