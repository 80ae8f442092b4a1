"""GPU tests of the continuous-batching server (include/xgpu_serve.h).

Each int8 stream's dumps must be bit-exact (int64) against the oracle summed
over the blocks of its integration, and identical whether the stream is
served alone or mixed with other streams of different shapes; FP16 streams'
dumps (double) must match the FP64 oracle within the fp32 accumulation bound.
"""
import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import workloads as W
from conftest import flush

pytestmark = pytest.mark.gpu

# (nstation, nfrequency, block ntime, blocks per dump); the last integrates
# 20 x 4096 samples, past the int32 bound, so its integration spills.
STREAMS = [(64, 8, 256, 4), (96, 5, 512, 3), (200, 3, 1024, 2), (128, 12, 256, 5), (8, 2, 4096, 20)]


def make_blocks(k, ns, nf, nt, count):
    blocks = []
    for b in range(count):
        hb = X.HostBuffer(nt * nf * ns * 4)
        hb.array[:] = oracle.fill_uniform(hb.nbytes, 1000 * k + b, -127, 127)
        blocks.append(hb)
    return blocks


def expected_dumps(ns, nf, nt, K, blocks):
    out = []
    for d in range(len(blocks) // K):
        acc = None
        for hb in blocks[d * K:(d + 1) * K]:
            v = oracle.xengine_i64(hb.array, ns, nf, nt)
            acc = v if acc is None else acc + v
        out.append(acc)
    return out


def serve(stream_ids, blocks_by_stream, interleave=True, **kw):
    srv = X.Server(**kw)
    opened = {}
    for k in stream_ids:
        ns, nf, nt, K = STREAMS[k]
        ndump = len(blocks_by_stream[k]) // K
        opened[k] = srv.open_stream(ns, nf, K, dump_slots=ndump)
    order = []
    if interleave:
        for i in range(max(len(blocks_by_stream[k]) for k in stream_ids)):
            for k in stream_ids:
                if i < len(blocks_by_stream[k]):
                    order.append((k, i))
    else:
        order = [(k, i) for k in stream_ids for i in range(len(blocks_by_stream[k]))]
    for k, i in order:
        srv.submit(opened[k][0], blocks_by_stream[k][i], STREAMS[k][2])
    srv.drain()
    result = {}
    for k in stream_ids:
        sid, ring = opened[k]
        ns, nf, nt, K = STREAMS[k]
        ndump = len(blocks_by_stream[k]) // K
        assert srv.dumps_done(sid) == ndump
        assert srv.blocks_done(sid) == len(blocks_by_stream[k])
        per = 2 * oracle.tri_len(ns, nf)
        result[k] = [ring.array[d * per:(d + 1) * per].copy() for d in range(ndump)]
    stats = srv.stats()
    srv.close()
    return result, stats


def test_serving_mix_is_exact_and_equals_alone():
    blocks = {k: make_blocks(k, *STREAMS[k][:3], 2 * STREAMS[k][3]) for k in range(len(STREAMS))}
    flush("[serve] mixed streams")
    mixed, stats = serve(list(range(len(STREAMS))), blocks)
    assert stats["blocks_done"] == sum(len(v) for v in blocks.values())
    assert stats["mean_batch"] > 1.0  # blocks of different streams shared launches
    for k, (ns, nf, nt, K) in enumerate(STREAMS):
        exp = expected_dumps(ns, nf, nt, K, blocks[k])
        for d, e in enumerate(exp):
            assert np.array_equal(mixed[k][d], e), (k, d)
    for k in (0, 2, 4):
        flush(f"[serve] stream {k} alone")
        alone, _ = serve([k], blocks, interleave=False)
        for d in range(len(alone[k])):
            assert np.array_equal(alone[k][d], mixed[k][d]), (k, d)
    for bl in blocks.values():
        for hb in bl:
            hb.free()


def test_serving_admission_under_small_staging_pool():
    """A pool that holds only ~2 blocks still serves everything, in order."""
    k = 1
    ns, nf, nt, K = STREAMS[k]
    blocks = {k: make_blocks(k, ns, nf, nt, 2 * K)}
    flush("[serve] small staging pool")
    res, stats = serve([k], blocks, staging_bytes=2 * nt * nf * ns * 4 + 512, max_batch=4)
    assert stats["staging_peak_bytes"] <= 2 * nt * nf * ns * 4 + 512
    exp = expected_dumps(ns, nf, nt, K, blocks[k])
    for d, e in enumerate(exp):
        assert np.array_equal(res[k][d], e)
    for hb in blocks[k]:
        hb.free()


# ---------------------------------------------------------------------------
# The serving mix as configured: 8 streams, NSTATION from {64, 96, 128, 200,
# 256, 352, 512}, NFREQUENCY in [8, 128], block NTIME from {256 .. 4096},
# K in [4, 64] blocks per dump, one stream integrating past the int32 bound.
# Each stream cycles through D = min(K, 3) distinct blocks (uniform int8 in
# [-127, 127]), so its integration is sum_b count_b V(block b): the oracle
# needs V of D blocks only.  Even streams submit host blocks (staged back to
# back and fused), odd streams submit blocks from a device ring of 2D slots
# (slot r holds block r % D; consecutive submissions are adjacent in memory
# except at the wrap, so they fuse too).
# ---------------------------------------------------------------------------
MIX = W.serving_mix(5)


class MixStream:
    def __init__(self, k, ns, nf, nt, K):
        self.k, self.ns, self.nf, self.nt, self.K = k, ns, nf, nt, K
        self.D = min(K, 3)
        self.nbytes = nt * nf * ns * 4
        self.device = bool(k & 1)
        self.host = []
        for b in range(self.D):
            hb = X.HostBuffer(self.nbytes)
            hb.array[:] = oracle.fill_uniform(self.nbytes, 7000 + 10 * k + b, -127, 127)
            self.host.append(hb)
        self.ring = None
        if self.device:
            self.ring = X.DeviceBuffer(2 * self.D * self.nbytes)
            for r in range(2 * self.D):
                X.check(X.core().xgpuAmdMemcpy(self.ring.ptr + r * self.nbytes,
                                               self.host[r % self.D].ptr, self.nbytes, 1))

    def submit(self, srv, sid, i):
        if self.device:
            srv.submit_device(sid, self.ring.ptr + (i % (2 * self.D)) * self.nbytes, self.nt)
        else:
            srv.submit(sid, self.host[i % self.D], self.nt)

    def counts(self):
        c = [0] * self.D
        for i in range(self.K):
            c[i % self.D] += 1
        return c

    def free(self):
        for hb in self.host:
            hb.free()
        if self.ring is not None:
            self.ring.free()


def serve_mix(streams, dump_device=False, **kw):
    """One dump per stream; submissions interleaved round-robin.  With
    dump_device the dump rings are device memory."""
    srv = X.Server(**kw)
    ids = {st.k: srv.open_stream(st.ns, st.nf, st.K, dump_slots=1, dump_device=dump_device)
           for st in streams}
    for i in range(max(st.K for st in streams)):
        for st in streams:
            if i < st.K:
                st.submit(srv, ids[st.k][0], i)
    srv.drain()
    out = {}
    for st in streams:
        sid, ring = ids[st.k]
        assert srv.dumps_done(sid) == 1 and srv.blocks_done(sid) == st.K, st.k
        n = 2 * oracle.tri_len(st.ns, st.nf)
        out[st.k] = ring.download(np.int64, n) if dump_device else ring.array[:n].copy()
    stats = srv.stats()
    srv.close()
    return out, stats


def check_stream_dump(st, dump):
    """All autocorrelations (from the input power) and every visibility of the
    first and last channel (exact int64 oracle) of one stream's dump."""
    ns, nf, nt = st.ns, st.nf, st.nt
    tri = dump.reshape(nf, ns * (ns + 1) // 2, 4, 2)
    counts = st.counts()
    power = np.zeros((nf, ns, 2), dtype=np.int64)
    chans = [0, nf - 1]
    ref = np.zeros((2, ns * (ns + 1) // 2, 4, 2), dtype=np.int64)
    for b, hb in enumerate(st.host):
        x = hb.array.reshape(nt, nf, ns, 2, 2).astype(np.int32)
        power += counts[b] * (x * x).sum(axis=(0, 4), dtype=np.int64)
        ref += counts[b] * oracle.xengine_i64_channels(hb.array, ns, nf, nt, chans).reshape(ref.shape)
    d = np.arange(ns) * (np.arange(ns) + 1) // 2 + np.arange(ns)
    assert np.array_equal(tri[:, d][:, :, [0, 3], 0], power), st.k
    assert not tri[:, d][:, :, [0, 3], 1].any(), st.k
    assert np.array_equal(tri[chans], ref), st.k


def test_serving_configured_mix_exact_and_equal_alone():
    assert any(nt * K > W.I32_SAMPLE_BOUND for _, _, nt, K in MIX)
    flush(f"[serve] configured mix {MIX}")
    streams = [MixStream(k, *MIX[k]) for k in range(len(MIX))]
    try:
        mixed, stats = serve_mix(streams, staging_bytes=4 << 30)
        assert stats["blocks_done"] == sum(st.K for st in streams)
        assert stats["mean_batch"] > 1.0    # several streams share launches
        for st in streams:
            flush(f"[serve] check stream {st.k} {MIX[st.k]} device={st.device}")
            check_stream_dump(st, mixed[st.k])
        spill = [st for st in streams if st.nt * st.K > W.I32_SAMPLE_BOUND][0]
        for st in {streams[0], streams[1], spill}:
            flush(f"[serve] stream {st.k} alone, dump ring in device memory")
            alone, _ = serve_mix([st], dump_device=True, staging_bytes=4 << 30)
            assert np.array_equal(alone[st.k], mixed[st.k]), st.k
    finally:
        for st in streams:
            st.free()


def test_serving_device_blocks_fuse_only_when_adjacent():
    """While a 128 MiB host block of another stream is staged and correlated
    (one launch in flight at most), a stream's K device blocks queue up: back
    to back in memory they fuse into one job, scattered they run as K jobs.
    The dumps are the same either way."""
    ns, nf, nt, K = 64, 4, 512, 6
    nbytes = nt * nf * ns * 4
    x = oracle.fill_uniform(K * nbytes, 77, -127, 127)
    ref = sum(oracle.xengine_i64(x[i * nbytes:(i + 1) * nbytes], ns, nf, nt) for i in range(K))
    flush("[serve] device blocks, adjacent and scattered")
    adj = X.DeviceBuffer.from_array(x)
    gap = X.DeviceBuffer(2 * K * nbytes)
    for i in range(K):
        gap.upload(x[i * nbytes:(i + 1) * nbytes], 2 * i * nbytes)
    bns, bnf, bnt = 512, 16, 4096
    blocker = X.HostBuffer(bnt * bnf * bns * 4)
    blocker.array[:] = 1
    for buf, stride, fused in ((adj, nbytes, True), (gap, 2 * nbytes, False)):
        srv = X.Server(max_inflight=1)
        bid, _ = srv.open_stream(bns, bnf, 1, dump_slots=1)
        sid, ring = srv.open_stream(ns, nf, K, dump_slots=1)
        srv.submit(bid, blocker, bnt)
        for i in range(K):
            srv.submit_device(sid, buf.ptr + i * stride, nt)
        srv.drain()
        st = srv.stats()
        assert np.array_equal(ring.array[: ref.size], ref), stride
        # scattered blocks are one job each (a job may still share a launch
        # with the blocker's); adjacent ones fuse
        if fused:
            assert st["mean_fused"] > 1.0, st
        else:
            assert st["mean_fused"] == 1.0, st
        srv.close()
    blocker.free()
    adj.free()
    gap.free()


def test_serving_device_dump_ring_slots():
    """Dumps into a device ring of 2 slots: dump k lands in slot k % 2, and
    dumps_done counts them; each dump is the integration of its own K blocks."""
    ns, nf, nt, K = 64, 4, 512, 2
    nbytes = nt * nf * ns * 4
    x = oracle.fill_uniform(3 * K * nbytes, 78, -127, 127)
    refs = [sum(oracle.xengine_i64(x[(d * K + i) * nbytes:(d * K + i + 1) * nbytes], ns, nf, nt)
                for i in range(K)) for d in range(3)]
    flush("[serve] device dump ring, 2 slots, 3 dumps")
    blocks = X.DeviceBuffer.from_array(x)
    srv = X.Server()
    sid, ring = srv.open_stream(ns, nf, K, dump_slots=2, dump_device=True)
    assert isinstance(ring, X.DeviceBuffer)
    n = refs[0].size
    for d in range(3):
        for i in range(K):
            srv.submit_device(sid, blocks.ptr + (d * K + i) * nbytes, nt)
        srv.drain()
        assert srv.dumps_done(sid) == d + 1
        assert np.array_equal(ring.download(np.int64, n, offset=(d % 2) * n * 8), refs[d]), d
    srv.close()
    blocks.free()


def test_serving_fp16_streams_beside_int8():
    """FP16 streams (host blocks; device blocks back to back, so they fuse)
    served beside an int8 stream: the FP16 dumps (double, triangular) match
    the FP64 oracle of the same blocks within the fp32 accumulation bound, the
    int8 dump is exact.  FP16 and int8 jobs never share a launch."""
    from test_gpu_parity import assert_f16_close, f16_input
    specs = [("f16", 64, 4, 256, 5, False), ("f16", 96, 3, 512, 4, True), ("i8", 64, 2, 256, 3, False)]
    flush(f"[serve] fp16 beside int8 {specs}")
    srv = X.Server()
    keep, ids = [], []
    for s, (kind, ns, nf, nt, K, dev) in enumerate(specs):
        comp = X.XGPU_FLOAT16 if kind == "f16" else X.XGPU_INT8
        sid, ring = srv.open_stream(ns, nf, K, compute=comp, dump_slots=1)
        if kind == "f16":
            blocks = [f16_input(nt, nf, ns, 900 + 10 * s + b) for b in range(K)]
        else:
            nb = nt * nf * ns * 4
            blocks = [oracle.fill_uniform(nb, 950 + b, -127, 127).reshape(nt, nf, ns, 2, 2)
                      for b in range(K)]
        if dev:
            d = X.DeviceBuffer.from_array(np.concatenate(blocks, axis=0))
            keep.append(d)
            for b in range(K):
                srv.submit_device(sid, d.ptr + b * blocks[0].nbytes, nt)
        else:
            hbs = []
            for blk in blocks:
                hb = X.HostBuffer(blk.nbytes)
                hb.array[:] = blk.reshape(-1).view(np.int8)
                hbs.append(hb)
            keep += hbs
            for hb in hbs:
                srv.submit(sid, hb, nt)
        ids.append((sid, ring, blocks))
    srv.drain()
    for (kind, ns, nf, nt, K, dev), (sid, ring, blocks) in zip(specs, ids):
        assert srv.dumps_done(sid) == 1
        x = np.concatenate(blocks, axis=0)
        n = 2 * oracle.tri_len(ns, nf)
        got = ring.array[:n]
        if kind == "f16":
            ref = oracle.xengine_f64_f16(x, ns, nf, K * nt)
            # one more fp32 rounding per job that accumulates: K at most
            assert_f16_close(got, ref, ns, nf, K * nt + 8 * K)
        else:
            assert np.array_equal(got, oracle.xengine_i64(x.reshape(-1), ns, nf, K * nt))
    srv.close()
    for b in keep:
        b.free()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_serving_randomized(seed):
    """Random stream shapes, host or device blocks (device blocks back to back,
    so they fuse), host or device dump rings of 2 slots, random staging pool,
    batch and in-flight limits; two threads submit in a random interleaving
    (each owns half the streams, so a stream's blocks stay in order).  Every
    stream's last two dumps are exact."""
    import threading
    rng = np.random.default_rng(seed)
    D = 3
    srv = X.Server(staging_bytes=int(rng.choice([4 << 20, 64 << 20])),
                   max_batch=int(rng.integers(1, 17)), max_inflight=int(rng.integers(1, 4)))
    streams = []
    for k in range(6):
        ns = int(rng.choice([8, 16, 64, 96, 200]))
        nf, nt, K = int(rng.integers(1, 6)), int(rng.choice([32, 100, 256, 512])), int(rng.integers(1, 7))
        dev, ddev = bool(rng.integers(2)), bool(rng.integers(2))
        nb = nt * nf * ns * 4
        x = oracle.fill_uniform(D * K * nb, 100 * seed + k, -127, 127)
        sid, ring = srv.open_stream(ns, nf, K, dump_slots=2, dump_device=ddev)
        src = X.DeviceBuffer.from_array(x) if dev else X.HostBuffer(x.nbytes)
        if not dev:
            src.array[:] = x
        streams.append(dict(ns=ns, nf=nf, nt=nt, K=K, dev=dev, ddev=ddev, nb=nb, x=x, sid=sid,
                            ring=ring, src=src))
    flush(f"[serve] randomized seed {seed}: "
          f"{[(s['ns'], s['nf'], s['nt'], s['K'], s['dev'], s['ddev']) for s in streams]}")
    errors = []

    def submitter(mine, r):
        try:
            nxt = {i: 0 for i in mine}
            while any(nxt[i] < D * streams[i]["K"] for i in mine):
                i = mine[int(r.integers(len(mine)))]
                s = streams[i]
                if nxt[i] == D * s["K"]:
                    continue
                b = nxt[i]
                nxt[i] += 1
                if s["dev"]:
                    srv.submit_device(s["sid"], s["src"].ptr + b * s["nb"], s["nt"])
                else:
                    srv.submit(s["sid"], s["src"].ptr + b * s["nb"], s["nt"])
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    th = [threading.Thread(target=submitter, args=(mine, np.random.default_rng(seed * 10 + t)))
          for t, mine in enumerate(([0, 1, 2], [3, 4, 5]))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    srv.drain()
    for s in streams:
        assert srv.dumps_done(s["sid"]) == D
        n = 2 * oracle.tri_len(s["ns"], s["nf"])
        ring = s["ring"].download(np.int64) if s["ddev"] else s["ring"].array.copy()
        for d in (D - 2, D - 1):  # dump d sits in slot d % 2
            ref = sum(oracle.xengine_i64(s["x"][(d * s["K"] + b) * s["nb"]:(d * s["K"] + b + 1) * s["nb"]],
                                         s["ns"], s["nf"], s["nt"]) for b in range(s["K"]))
            assert np.array_equal(ring[(d % 2) * n:(d % 2 + 1) * n], ref), (seed, s["ns"], d)
    srv.close()
    for s in streams:
        s["src"].free()
