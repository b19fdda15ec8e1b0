"""GPU parity of the batched launch and the serving loop against the oracle and
the golden vectors of the compiled reference."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
DT = {0: np.float64, 1: np.float64, 2: np.uint32}


def say(*a):
    print("[case]", *a, flush=True)


class Packed:
    """Segments laid out back to back in one device buffer (element-aligned only)."""

    def __init__(self, counts, fmts, gap_elems=1):
        from mrg8_amd.hip import DeviceBuffer

        self.offsets, pos = [], 0
        for c, f in zip(counts, fmts):
            esz = np.dtype(DT[f]).itemsize
            pos = (pos + esz - 1) // esz * esz
            self.offsets.append(pos)
            pos += (int(c) + gap_elems) * esz
        self.buf = DeviceBuffer(max(pos, 8))
        self.buf.fill_byte(0xAB)
        self.counts, self.fmts = counts, fmts

    def ptr(self, i):
        return self.buf.ptr + self.offsets[i]

    def fetch(self, i):
        return self.buf.copy_to_host(DT[self.fmts[i]], int(self.counts[i]), self.offsets[i])


@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
@pytest.mark.parametrize("shapes", [((256, 4, 4096), (512, 2, 16384)), ((1024, 0, 4096), (256, 0, 512)),
                                    ((512, -2, 2048), (256, -6, 4096))],
                         ids=["persistent", "tile-per-wave", "address-order"])
def test_batch_mixed_segments(gpu, oracle_mod, shapes, graph):
    rng = np.random.default_rng(21)
    n = 200
    states = rng.integers(0, 2**31, size=(n, 8)).astype(np.uint32)
    states[::17, rng.integers(0, 8)] = 0
    counts = rng.integers(1, 70000, size=n)
    counts[::23] = 0
    fmts = rng.integers(0, 3, size=n)
    offs = [int(x) for x in rng.integers(0, 2**63, size=n, dtype=np.uint64)]
    offs[:20] = list(range(20))  # small offsets take the shift rows of the jump
    pk = Packed(counts, fmts, gap_elems=int(rng.integers(1, 4)))
    b = gpu.Batch(max_streams=n, max_tiles=1 << 20)
    b.use_graph(graph)
    b.set_config(*shapes[0])
    segs = [(states[i], offs[i], int(counts[i]), pk.ptr(i), int(fmts[i])) for i in range(n)]
    b.set(segs)
    for rep in range(3):  # repeated runs (graph replays) give identical results
        say("batch-mixed", shapes, graph, rep)
        b.run()
        gpu.hip.synchronize()
        for i in range(n):
            if counts[i] == 0:
                continue
            want, _ = oracle_mod.generate(oracle_mod.jump(states[i], offs[i]), int(counts[i]), int(fmts[i]))
            got = pk.fetch(i)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (rep, i)
        b.set_config(*shapes[1])  # a new shape, same outputs


@pytest.mark.parametrize("shape", [None, (256, -2, 2048)], ids=["default", "address-order"])
@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
def test_batch_restaged_on_another_stream(gpu, oracle_mod, graph, shape):
    """A batch run on one stream, re-staged and run at once on another: the
    new block's copy waits for the first run, and both runs' outputs hold.
    The two stagings have different tile counts: a persistent grid's graph
    is reused for the second (its kernel reads the count from the block)."""
    from mrg8_amd import hip

    rng = np.random.default_rng(5)
    n = 64
    st = rng.integers(1, 2**31 - 1, size=(2, n, 8)).astype(np.uint32)
    cnt = rng.integers(1000, 300000, size=(2, n))
    offs = rng.integers(0, 2**40, size=(2, n), dtype=np.uint64)
    bufs = [hip.DeviceBuffer(int(cnt[r].sum()) * 8) for r in range(2)]
    starts = [np.concatenate([[0], np.cumsum(cnt[r])[:-1]]) for r in range(2)]
    a_s, b_s = hip.Stream(), hip.Stream()
    b = gpu.Batch(max_streams=n, max_tiles=1 << 16)
    if shape:
        b.set_config(*shape)
    b.use_graph(graph)
    say("restaged", graph, shape)
    for r, s in ((0, a_s), (1, b_s)):
        b.set([(st[r, i], int(offs[r, i]), int(cnt[r, i]), bufs[r].ptr + 8 * int(starts[r][i]), 0) for i in range(n)])
        b.run(s)
    b_s.synchronize()
    a_s.synchronize()
    for r in range(2):
        got = bufs[r].copy_to_host(np.float64, int(cnt[r].sum()))
        for i in range(0, n, 7):
            want, _ = oracle_mod.generate(oracle_mod.jump(st[r, i], int(offs[r, i])), int(cnt[r, i]), 0)
            part = got[int(starts[r][i]) : int(starts[r][i]) + int(cnt[r, i])]
            assert np.array_equal(part.view(np.uint64), want.view(np.uint64)), (r, i)
    b.close()
    for x in bufs:
        x.free()
    a_s.destroy()
    b_s.destroy()


@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
def test_address_order_runs_on_several_streams(gpu, oracle_mod, graph):
    """An address-order grid (blocks_per_cu < 0) takes its tiles from counters
    in the batch's device block, which its last block sets back to zero: runs
    of one unchanged batch issued on several streams without waiting, and a
    tile-per-wave run between them, are serialised by the batch, so a later
    run into a cleared buffer still writes every output."""
    from mrg8_amd import hip

    rng = np.random.default_rng(9)
    n = 96
    st = rng.integers(1, 2**31 - 1, size=(n, 8)).astype(np.uint32)
    cnt = rng.integers(1000, 200000, size=n)
    offs = rng.integers(0, 2**62, size=n, dtype=np.uint64)
    total = int(cnt.sum())
    starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    buf = hip.DeviceBuffer(total * 8)
    streams = [hip.Stream() for _ in range(3)]
    b = gpu.Batch(max_streams=n, max_tiles=1 << 16)
    b.use_graph(graph)
    b.set_config(256, -2, 2048)
    b.set([(st[i], int(offs[i]), int(cnt[i]), buf.ptr + 8 * int(starts[i]), 0) for i in range(n)])
    want = [oracle_mod.generate(oracle_mod.jump(st[i], int(offs[i])), int(cnt[i]), 0)[0] for i in range(0, n, 5)]
    for rnd, order in enumerate(((0, 1, 2, 0), (2, 1, 1, 0))):
        say("address-order streams", graph, rnd)
        for k in order:
            b.run(streams[k])
        b.set_config(512, 0, 2048)  # a tile-per-wave run in between, on another stream
        b.run(streams[(order[-1] + 1) % 3])
        b.set_config(256, -2, 2048)
        b.run(streams[order[-1]])
        for s in streams:
            s.synchronize()
        buf.fill_byte(0)
        b.run(streams[1])
        streams[1].synchronize()
        got = buf.copy_to_host(np.float64, total)
        for j, i in enumerate(range(0, n, 5)):
            part = got[int(starts[i]):int(starts[i]) + int(cnt[i])]
            assert np.array_equal(part.view(np.uint64), want[j].view(np.uint64)), (rnd, i)
    b.close()
    buf.free()
    for s in streams:
        s.destroy()


@pytest.mark.parametrize("bpc", [1, 0, -2], ids=["persistent", "tile-per-wave", "address-order"])
def test_batch_tile_heads(gpu, oracle_mod, bpc):
    """Segments at many element offsets within a tile-alignment unit (the
    largest power of two dividing tile * element size), with counts around the
    head (the outputs before the first such boundary, which tile 0 of a
    segment also writes) and the tile size: outputs match the oracle and the
    bytes around every segment stay untouched."""
    from mrg8_amd.hip import DeviceBuffer

    tile = 256
    cases = []
    for fmt, esz in ((0, 8), (2, 4), (1, 8)):
        align = tile * esz  # a power of two here
        for m in sorted(set(range(0, 128, esz)) | {align // 2, align // 2 + esz, align - 2 * esz, align - esz}):
            head = ((align - m) % align) // esz
            for c in sorted({1, max(1, head - 1), max(1, head), head + 1, head + tile, head + tile + 1,
                             head + 3 * tile + 5}):
                cases.append((fmt, esz, m, c))
    slot = (max(m + c * esz for _, esz, m, c in cases) + 256 + 2047) // 2048 * 2048
    rng = np.random.default_rng(77)
    states = rng.integers(1, 2**31 - 1, size=(len(cases), 8)).astype(np.uint32)
    offs = [int(x) for x in rng.integers(0, 2**63, size=len(cases), dtype=np.uint64)]
    buf = DeviceBuffer(slot * len(cases) + 2048)
    buf.fill_byte(0xAB)
    base = (buf.ptr + 2047) // 2048 * 2048  # every slot starts on a 2 KiB boundary
    b = gpu.Batch(max_streams=len(cases), max_tiles=1 << 16)
    b.set_config(256, bpc, tile)  # persistent: a small grid, every wave takes several tiles
    b.set([(states[i], offs[i], c, base + i * slot + m, fmt) for i, (fmt, _, m, c) in enumerate(cases)])
    say("batch-tile-heads", len(cases))
    b.run()
    gpu.hip.synchronize()
    host = buf.copy_to_host(np.uint8, slot * len(cases), base - buf.ptr)
    for i, (fmt, esz, m, c) in enumerate(cases):
        want, _ = oracle_mod.generate(oracle_mod.jump(states[i], offs[i]), c, fmt)
        region = host[i * slot:(i + 1) * slot]
        assert region[m:m + c * esz].tobytes() == want.tobytes(), (fmt, m, c)
        assert (region[:m] == 0xAB).all() and (region[m + c * esz:] == 0xAB).all(), (fmt, m, c)


def test_batch_golden_set_state_streams(gpu, oracle_mod):
    rows = GOLD["set_state_streams"]
    counts = [r["n"] for r in rows]
    fmts = [r["fmt"] for r in rows]
    pk = Packed(counts, fmts)
    b = gpu.Batch(max_streams=len(rows))
    b.set([(r["init"], 0, r["n"], pk.ptr(i), r["fmt"]) for i, r in enumerate(rows)])
    say("batch-golden")
    b.run()
    gpu.hip.synchronize()
    for i, r in enumerate(rows):
        assert f"{oracle_mod.fnv1a(pk.fetch(i)):016x}" == r["fnv1a64"], i


def test_batch_golden_jump_then_randint(gpu, oracle_mod):
    """The jump-heavy shape (63-bit steps from one seed, 256 words each), pinned
    directly against the compiled reference's outputs."""
    rows = GOLD["jump_then_randint"]
    s0 = gpu.seed_state(GOLD["seed"])
    pk = Packed([256] * len(rows), [2] * len(rows))
    b = gpu.Batch(max_streams=len(rows))
    b.set_config(256, 8, 256)
    b.set([(s0, int(r["step"]), 256, pk.ptr(i), 2) for i, r in enumerate(rows)])
    say("batch-jumps")
    b.run()
    gpu.hip.synchronize()
    for i, r in enumerate(rows):
        got = pk.fetch(i)
        assert f"{oracle_mod.fnv1a(got):016x}" == r["fnv1a64"], i
        assert [str(x) for x in got[:4]] == r["first"]


def test_batch_autotune_keeps_outputs(gpu, oracle_mod, tmp_path):
    rng = np.random.default_rng(5)
    n = 64
    states = rng.integers(1, 2**31 - 1, size=(n, 8)).astype(np.uint32)
    counts = rng.integers(1000, 300000, size=n)
    pk = Packed(counts, [0] * n)
    b = gpu.Batch(max_streams=n)
    b.set([(states[i], 0, int(counts[i]), pk.ptr(i), 0) for i in range(n)])
    say("autotune")
    cache = str(tmp_path / "tune.txt")
    cfg = b.autotune(cache_path=cache)
    assert cfg["tile_outputs"] in (2048, 4096, 8192, 16384, 65536)
    assert os.path.exists(cache)
    cfg2 = b.autotune(cache_path=cache)  # second call is served from the cache
    assert cfg2 == cfg
    pk.buf.fill_byte(0)
    b.run()
    gpu.hip.synchronize()
    sums = oracle_mod.stream_wsums(states, np.zeros(n, dtype=np.uint64), counts, np.zeros(n, dtype=np.int32))
    for i in range(n):
        assert oracle_mod.wsum(pk.fetch(i)) == int(sums[i]), i


def test_batch_errors(gpu):
    from mrg8_amd.capi import Mrg8Error

    b = gpu.Batch(max_streams=2, max_tiles=4)
    with pytest.raises(Mrg8Error):
        b.set([([1] * 8, 0, 10, 0, 0)])  # null output
    with pytest.raises(Mrg8Error):
        b.set([([1] * 8, 0, 10, 8, 7)])  # bad format
    with pytest.raises(Mrg8Error):
        b.set([([2**31] + [1] * 7, 0, 10, 8, 0)])  # word above p
    with pytest.raises(Mrg8Error):
        b.set([([1] * 8, 0, 10, 8, 0)] * 3)  # too many segments
    with pytest.raises(Mrg8Error):
        b.set([([1] * 8, 0, 10**9, 8, 0)])  # too many tiles
    with pytest.raises(Mrg8Error):
        b.set_config(100, 1, 1024)
    with pytest.raises(Mrg8Error):
        b.set_config(256, 0, 1 << 41)  # tile above 2^40
    with pytest.raises(Mrg8Error):
        b.set_config(256, 0, 2**64 - 256)


@pytest.mark.parametrize("shape,graph", [(None, False), ((1024, 0, 4096), False), ((256, 4, 4096), True),
                                         ((1024, 0, 2048), True), ((512, -2, 2048), False), ((512, -2, 2048), True)],
                         ids=["default", "tile-per-wave", "persistent-graph", "tile-per-wave-graph", "address-order",
                              "address-order-graph"])
def test_c3_batched_full_size(gpu, oracle_mod, shape, graph):
    """4096 streams of log-uniform length in [2^10, 2^22), float64, one launch
    (a plain launch, or a replay of the captured graph, replayed twice with the
    output cleared in between); every stream checked by its position-weighted
    checksum."""
    from mrg8_amd import workloads
    from mrg8_amd.hip import DeviceBuffer

    states, lengths = workloads.batched()
    total = int(lengths.sum())
    say("c3", total, shape, "graph" if graph else "direct")
    buf = DeviceBuffer(total * 8)
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    b = gpu.Batch(max_streams=len(lengths))
    if shape:
        b.set_config(*shape)
    b.use_graph(graph)
    b.set([(states[i], 0, int(lengths[i]), buf.ptr + 8 * int(starts[i]), 0) for i in range(len(lengths))])
    want = oracle_mod.stream_wsums(states, np.zeros(len(lengths), dtype=np.uint64), lengths,
                                   np.zeros(len(lengths), dtype=np.int32))
    for replay in range(2 if graph or (shape and shape[1] < 0) else 1):
        if replay:
            buf.fill_byte(0)  # the second replay must write every output again
        b.run()
        gpu.hip.synchronize()
        host = buf.copy_to_host(np.float64, total)
        bad = [i for i in range(len(lengths))
               if oracle_mod.wsum(host[int(starts[i]):int(starts[i] + lengths[i])]) != int(want[i])]
        assert not bad, (replay, bad[:10])
        del host
    buf.free()


@pytest.mark.parametrize("shape", [(256, 8, 256), (1024, 0, 256)], ids=["persistent", "tile-per-wave"])
def test_c4_jump_heavy_full_size(gpu, oracle_mod, shape):
    """65536 streams from one seed, each jumped by a uniform step in [0, 2^63),
    then 256 words: every stream against the oracle's jump + generation."""
    from mrg8_amd import workloads
    from mrg8_amd.hip import DeviceBuffer

    steps = workloads.jump_steps()
    n = len(steps)
    s0 = gpu.seed_state(workloads.SEED)
    buf = DeviceBuffer(n * 1024)
    b = gpu.Batch(max_streams=n, max_tiles=n)
    b.set_config(*shape)
    b.set([(s0, int(steps[i]), 256, buf.ptr + 1024 * i, 2) for i in range(n)])
    say("c4", n, shape)
    b.run()
    gpu.hip.synchronize()
    host = buf.copy_to_host(np.uint32, n * 256).reshape(n, 256)
    buf.free()
    want = oracle_mod.stream_wsums(np.tile(s0, (n, 1)), steps, np.full(n, 256, dtype=np.uint64),
                                   np.full(n, 2, dtype=np.int32))
    weights = np.arange(256, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    got = np.sum(host.astype(np.uint64) * weights, axis=1, dtype=np.uint64)
    bad = np.flatnonzero(got != want).tolist()
    assert not bad, bad[:10]
    # spot-check full words for a few streams
    for i in (0, 1, n // 2, n - 1):
        w, _ = oracle_mod.generate(oracle_mod.jump(s0, int(steps[i])), 256, 2)
        assert np.array_equal(host[i], w)


# ---- serving -------------------------------------------------------------------------


def drain(srv, expect, on_done, timeout_s=60):
    """Progress the server until `expect` requests completed; each completion is
    passed to on_done and then released (admission is by output memory, so a
    caller that never releases stalls the server by design)."""
    import time

    t0, n = time.time(), 0
    while n < expect:
        for c in srv.progress(max_done=256, wait_ms=50):
            on_done(c)
            srv.release(c["request_id"])
            n += 1
        assert time.time() - t0 < timeout_s, f"server stalled after {n}/{expect} completions: {srv.stats()}"


def check_request(gpu, oracle_mod, c, base_state, fmt):
    from mrg8_amd import hip

    dt = DT[fmt]
    host = np.empty(int(c["count"]), dtype=dt)
    hip.check(hip._load().hipMemcpy(host.ctypes.data, hip.ctypes.c_void_p(c["out"]), host.nbytes, 2), "D2H")
    want, _ = oracle_mod.generate(oracle_mod.jump(base_state, int(c["start"])), int(c["count"]), fmt)
    return np.array_equal(host.view(np.uint8), want.view(np.uint8)), host


@pytest.mark.parametrize("grid", [{}, {"block_threads": 256, "blocks_per_cu": -2}], ids=["default", "address-order"])
def test_server_requests_match_oracle(gpu, oracle_mod, grid):
    """Every request's outputs equal the oracle's, with the serving loop's
    default persistent grid and with the address-order grid (its steps run on
    several streams, each on a batch block of its own)."""
    rng = np.random.default_rng(31)
    srv = gpu.Server(num_streams=64, max_output_bytes=64 << 20, base_seed=1000, step_outputs=1 << 20,
                     tile_outputs=8192, **grid)
    reqs = {}
    for k in range(300):
        sid = int(rng.integers(0, 64))
        cnt = int(rng.integers(1, 200000))
        fmt = int(rng.integers(0, 3))
        reqs[srv.submit(sid, cnt, fmt)] = (sid, fmt)
    say("server-oracle", grid)

    def check(c):
        sid, fmt = reqs[c["request_id"]]
        ok, _ = check_request(gpu, oracle_mod, c, oracle_mod.seed_state(1000 + sid), fmt)
        assert ok, (c["request_id"], sid, c["start"], c["count"])

    drain(srv, len(reqs), check)
    st = srv.stats()
    assert st["completed"] == 300 and st["bytes_inflight"] == 0 and st["queued"] == 0


def test_server_set_streams_bulk(gpu, oracle_mod):
    """Streams set in one upload (with start positions) serve from those
    states and positions; the others keep their seeded states."""
    rng = np.random.default_rng(41)
    states = rng.integers(1, 2**31 - 1, size=(20, 8)).astype(np.uint32)
    positions = [int(x) for x in rng.integers(0, 2**62, size=20, dtype=np.uint64)]
    srv = gpu.Server(num_streams=32, max_output_bytes=16 << 20, base_seed=500, step_outputs=1 << 18)
    srv.set_streams(states, positions, first=4)
    for sid in (3, 4, 13, 23, 24):
        st, pos = srv.get_stream(sid)
        if 4 <= sid < 24:
            assert list(st) == list(states[sid - 4]) and pos == positions[sid - 4]
        else:
            assert list(st) == list(oracle_mod.seed_state(500 + sid)) and pos == 0
    reqs = {}
    for sid in (3, 4, 13, 23, 24):
        reqs[srv.submit(sid, 3001, 0)] = sid
    say("server-set-streams")

    def check(c):
        sid = reqs[c["request_id"]]
        base = states[sid - 4] if 4 <= sid < 24 else oracle_mod.seed_state(500 + sid)
        ok, _ = check_request(gpu, oracle_mod, c, base, 0)
        assert ok, sid

    drain(srv, len(reqs), check)
    with pytest.raises(Exception):
        srv.set_streams(states, first=20)  # past the last stream


def test_server_alone_equals_mixed(gpu, oracle_mod):
    """A request gives the same outputs alone and in the middle of a mix."""
    seq = [(5, 1000, 0), (5, 5000, 2), (5, 77777, 0)]  # the third is the probe request
    # alone: one request at a time
    a = gpu.Server(num_streams=16, max_output_bytes=8 << 20, base_seed=7)
    outs_alone = []
    for sid, cnt, fmt in seq:
        a.submit(sid, cnt, fmt)

        def keep(c, sid=sid, fmt=fmt):
            ok, host = check_request(gpu, oracle_mod, c, oracle_mod.seed_state(7 + sid), fmt)
            assert ok
            outs_alone.append(host)

        drain(a, 1, keep)
    # mixed: the same requests on stream 5, interleaved with 200 others on other streams
    m = gpu.Server(num_streams=16, max_output_bytes=32 << 20, base_seed=7, step_outputs=1 << 18,
                   tile_outputs=4096)
    rng = np.random.default_rng(3)
    mine = {}
    for k in range(200):
        if k in (10, 80, 150):
            sid, cnt, fmt = seq[len(mine)]
            mine[m.submit(sid, cnt, fmt)] = len(mine)
        m.submit(int(rng.integers(6, 16)), int(rng.integers(1, 100000)), int(rng.integers(0, 3)))
    say("server-mix")
    outs_mixed = {}

    def grab(c):
        if c["request_id"] in mine:
            j = mine[c["request_id"]]
            sid, cnt, fmt = seq[j]
            ok, host = check_request(gpu, oracle_mod, c, oracle_mod.seed_state(7 + sid), fmt)
            assert ok
            outs_mixed[j] = host

    drain(m, 203, grab)
    for j in range(3):
        assert np.array_equal(outs_mixed[j].view(np.uint8), outs_alone[j].view(np.uint8)), j


def test_server_admission_by_memory(gpu, oracle_mod):
    cap = 4 << 20  # 4 MiB of output: at most four 1 MiB requests at once
    srv = gpu.Server(num_streams=8, max_output_bytes=cap, base_seed=3, step_outputs=1 << 16, tile_outputs=4096)
    ids = [srv.submit(i % 8, 131072, 0) for i in range(32)]  # 1 MiB each
    seen_max = 0
    done = {}
    say("server-admission")
    while len(done) < len(ids):
        got = srv.progress(max_done=4, wait_ms=20)
        st = srv.stats()  # completed-but-unreleased requests still hold their bytes
        seen_max = max(seen_max, st["bytes_inflight"])
        assert st["bytes_inflight"] <= cap
        assert st["inflight"] + len(got) <= 4  # memory admits at most four at a time
        for c in got:
            done[c["request_id"]] = c
            ok, _ = check_request(gpu, oracle_mod, c, oracle_mod.seed_state(3 + c["stream_id"]), 0)
            assert ok
            srv.release(c["request_id"])
    assert seen_max >= 1 << 20
    assert srv.stats()["bytes_inflight"] == 0


def test_server_errors_and_cancel(gpu):
    from mrg8_amd.capi import Mrg8Error

    srv = gpu.Server(num_streams=4, max_output_bytes=1 << 20, base_seed=1)
    with pytest.raises(Mrg8Error):
        srv.submit(4, 10, 0)  # unknown stream
    with pytest.raises(Mrg8Error):
        srv.submit(0, 1 << 20, 0)  # larger than the cap
    with pytest.raises(Mrg8Error):
        srv.release(12345)
    rid = srv.submit(1, 100, 2)  # queued: never progressed, so it can be cancelled
    srv.release(rid)
    st, pos = srv.get_stream(1)
    assert pos == 100  # the cancelled request's positions stay consumed
    srv.set_stream(1, [5] * 8, 0)
    assert srv.get_stream(1)[0].tolist() == [5] * 8


def test_server_replay_small_trace(gpu, oracle_mod):
    srv = gpu.Server(num_streams=32, max_output_bytes=64 << 20, base_seed=11, step_outputs=1 << 20)
    rng = np.random.default_rng(8)
    n = 2000
    arrivals = np.cumsum(rng.exponential(1e-4, size=n))
    streams = rng.integers(0, 32, size=n)
    sizes = np.floor(2.0 ** (8 + 10 * rng.random(n))).astype(np.uint64)
    say("replay")
    stats, lat = srv.replay(arrivals, streams, sizes, fmt=0, want_latencies=True)
    assert stats["completed"] == n and stats["outputs"] == int(sizes.sum())
    assert 0 < stats["p50_s"] <= stats["p99_s"] <= stats["max_s"]
    assert (lat > 0).all()
    # the replay consumed the streams' positions in trace order
    for s in range(32):
        assert srv.get_stream(s)[1] == int(sizes[streams == s].sum())


def test_server_cap_granule_and_blocked_queue(gpu, oracle_mod):
    """The cap counts whole 256-byte granules (a 1000-byte cap serves 768
    bytes), and a queue head that only a release can admit is reported as
    MRG8_ERR_BLOCKED instead of a silent stall."""
    from mrg8_amd.capi import ERR_BLOCKED, ERR_CAPACITY, Mrg8Error

    say("server-granule-blocked")
    srv = gpu.Server(num_streams=2, max_output_bytes=1000, base_seed=1)
    with pytest.raises(Mrg8Error) as e:
        srv.submit(0, 97, 0)  # 776 bytes take 1024 in granules
    assert e.value.code == ERR_CAPACITY
    srv.submit(0, 96, 0)  # 768 bytes fit exactly
    drain(srv, 1, lambda c: None)
    srv.close()

    srv = gpu.Server(num_streams=2, max_output_bytes=4096, base_seed=1)
    a = srv.submit(0, 300, 0)  # 2560 bytes in granules
    got = []
    while not got:
        got = srv.progress(max_done=4, wait_ms=50)
    assert got[0]["request_id"] == a  # completed, not released: holds its 2560 bytes
    b = srv.submit(1, 300, 0)
    with pytest.raises(Mrg8Error) as e:
        srv.progress(max_done=4, wait_ms=50)
    assert e.value.code == ERR_BLOCKED
    srv.release(a)
    done = []
    while not done:
        done = srv.progress(max_done=4, wait_ms=50)
    assert done[0]["request_id"] == b
    ok, _ = check_request(gpu, oracle_mod, done[0], oracle_mod.seed_state(2), 0)
    assert ok
    srv.release(b)
    srv.close()
