"""GPU tests of the continuous-batching server (include/xgpu_serve.h).

Each stream's dumps must be bit-exact (int64) against the oracle summed over
the blocks of its integration, and identical whether the stream is served
alone or mixed with other streams of different shapes.
"""
import numpy as np
import pytest

import oracle
import xgpu_amd as X
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
