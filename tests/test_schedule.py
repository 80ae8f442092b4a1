"""CPU tests of the persistent-kernel tile schedule (no GPU).

The kernels decode their work from the launch descriptor alone: XCD `x`'s
k-th tile is (job, channel, row block, column block) by `xg_tile` in
xg_kernels.hip, which the library also exports to the host as
`xgh_schedule`.  These tests check that every launch shape covers every tile
of every job's lower triangle exactly once (the property every visibility of
the parity tests relies on), and that the order keeps concurrent tiles on
shared input: at the headline shape (512 stations, 1024 inputs = 8 blocks,
36 tiles per channel, 32 workgroups per XCD) a round of 32 tiles works on one
channel, and the tiles left over per channel are packed into few input chunks.
"""
import ctypes as C
import itertools

import pytest

import xgpu_amd as X

XG_MAX_JOBS, XG_NXCD, XG_TILE = 16, 8, 128


class XgJob(C.Structure):
    _fields_ = [("input", C.c_void_p), ("out", C.c_void_p), ("tstride", C.c_longlong),
                ("plane", C.c_longlong), ("nstation", C.c_int), ("nfreq", C.c_int),
                ("ntime", C.c_int), ("nblk", C.c_int), ("tiles_per_chan", C.c_int),
                ("order", C.c_int), ("accumulate", C.c_int), ("out_kind", C.c_int),
                ("chan_base", C.c_int), ("pair_tiles", C.c_int)]


class XgLaunch(C.Structure):
    _fields_ = [("jobs", XgJob * XG_MAX_JOBS), ("njobs", C.c_int), ("wg_per_xcd", C.c_short),
                ("prefetch", C.c_char), ("stage", C.c_char),
                ("xcd_start", (C.c_int * (XG_MAX_JOBS + 1)) * XG_NXCD)]


def test_descriptor_sizes_match_the_c_layout():
    assert C.sizeof(XgJob) == 72
    assert C.sizeof(XgLaunch) == 72 * XG_MAX_JOBS + 8 + 4 * XG_NXCD * (XG_MAX_JOBS + 1)


def plan(jobs, cus=256, wg_per_cu=1):
    """jobs: list of (nstation, nfreq). Mirrors the engine's job fill."""
    L = XgLaunch()
    chan = 0
    for i, (ns, nf) in enumerate(jobs):
        J = L.jobs[i]
        J.nstation, J.nfreq, J.ntime = ns, nf, 64
        J.nblk = (2 * ns + XG_TILE - 1) // XG_TILE
        J.tiles_per_chan = J.nblk * (J.nblk + 1) // 2
        J.chan_base = chan
        chan += nf
    L.njobs = len(jobs)
    lib = X.core()
    lib.xg_plan_launch.argtypes = [C.POINTER(XgLaunch), C.c_int, C.c_int]
    lib.xg_plan_launch(C.byref(L), cus, wg_per_cu)
    return L


def schedule(L):
    """{xcd: [(job, f, rb, cb), ...]} in the order each XCD takes them."""
    lib = X.core()
    fn = lib.xgh_schedule
    fn.argtypes = [C.POINTER(XgLaunch), C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4
    out = {}
    for x in range(XG_NXCD):
        seq, k = [], 0
        v = [C.c_int() for _ in range(4)]
        while fn(C.byref(L), x, k, *[C.byref(a) for a in v]):
            seq.append(tuple(a.value for a in v))
            k += 1
        out[x] = seq
    return out


@pytest.mark.parametrize("jobs,cus,wg", [
    ([(512, 64)], 256, 1),            # headline shape, real-Gram kernel (1 WG per CU)
    ([(512, 64)], 256, 2),            # complex-plane kernel (2 WGs per CU)
    ([(2048, 16)], 256, 1),           # 4096 inputs: 32 blocks, 528 tiles per channel
    ([(256, 10)], 256, 2),
    ([(2, 1)], 256, 1),               # one tile in all
    ([(132, 3)], 256, 1),             # partial last block
    ([(64, 3), (96, 2), (200, 1), (16, 5)], 256, 1),   # ragged batch
    ([(512, 5), (1000, 3)], 40, 2),   # few CUs: many rounds per XCD
    ([(16 + 12 * j, 1 + j % 5) for j in range(16)], 256, 1),   # a batch of the maximum 16 jobs
    ([(512, 8191)], 256, 1),          # the largest real-Gram block: 294,876 tiles
])
def test_every_tile_exactly_once(jobs, cus, wg):
    L = plan(jobs, cus, wg)
    sched = schedule(L)
    seen = [t for x in range(XG_NXCD) for t in sched[x]]
    expect = []
    for j, (ns, nf) in enumerate(jobs):
        nb = (2 * ns + XG_TILE - 1) // XG_TILE
        expect += [(j, f, rb, cb) for f in range(nf) for rb in range(nb) for cb in range(rb + 1)]
    assert sorted(seen) == sorted(expect)
    assert len(seen) == len(set(seen))
    # a channel's tiles all sit on one XCD: f + chan_base = xcd (mod 8)
    base = [sum(nf for _, nf in jobs[:j]) for j in range(len(jobs))]
    for x in range(XG_NXCD):
        assert all((f + base[j]) % XG_NXCD == x for j, f, _, _ in sched[x])


def test_tile_order_is_a_permutation_of_the_triangle():
    lib = X.core()
    L = plan([(2, 1)])
    fn = lib.xgh_schedule
    fn.argtypes = [C.POINTER(XgLaunch), C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4
    for nblk in list(range(1, 41)) + [64, 100]:
        # one job, one channel, W = 1: k walks the channel order itself
        L = plan([(nblk * XG_TILE // 2, 1)], cus=8, wg_per_cu=1)
        L.wg_per_xcd = 1
        v = [C.c_int() for _ in range(4)]
        got = []
        k = 0
        while fn(C.byref(L), 0, k, *[C.byref(a) for a in v]):
            got.append((v[2].value, v[3].value))
            k += 1
        assert sorted(got) == [(r, c) for r in range(nblk) for c in range(r + 1)], nblk


def pair_count(nblk):
    return sum(nblk - 2 * cp for cp in range((nblk + 1) // 2))


def test_pair_tile_order_covers_the_triangle():
    """Pair tiles (rb, cp) of the K-interleaved kernel: 128 row inputs by 256
    column inputs with 2 cp <= rb.  Every pair is taken once, and every input
    pair (m, n) with n <= m + 1 (the station-level lower triangle) lies in
    exactly one of them."""
    lib = X.core()
    fn = lib.xgh_schedule
    fn.argtypes = [C.POINTER(XgLaunch), C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4
    for nblk in list(range(1, 41)) + [64]:
        L = plan([(nblk * XG_TILE // 2, 1)], cus=8, wg_per_cu=1)
        L.jobs[0].pair_tiles = 1
        L.jobs[0].tiles_per_chan = pair_count(nblk)
        lib.xg_plan_launch(C.byref(L), 8, 1)
        L.wg_per_xcd = 1
        v = [C.c_int() for _ in range(4)]
        got = []
        k = 0
        while fn(C.byref(L), 0, k, *[C.byref(a) for a in v]):
            got.append((v[2].value, v[3].value))
            k += 1
        assert sorted(got) == sorted((r, c) for r in range(nblk) for c in range(r // 2 + 1)), nblk
        if nblk <= 8:
            ninput = nblk * XG_TILE
            cover = {}
            for rb, cp in got:
                for m in range(rb * XG_TILE, (rb + 1) * XG_TILE):
                    for n in range(2 * cp * XG_TILE, min((2 * cp + 2) * XG_TILE, m + 2)):
                        cover[(m, n)] = cover.get((m, n), 0) + 1
            need = {(m, n) for m in range(ninput) for n in range(min(ninput, m + 2))
                    if (n >> 1) <= (m >> 1)}
            assert need <= set(cover) and all(cover[p] == 1 for p in need), nblk


def chunk_streams(sched, W):
    """Input chunk-streams fetched from memory if each round of W concurrent
    tiles fetches every (job, channel, chunk) it touches once."""
    total = 0
    for x, seq in sched.items():
        for r in range(0, len(seq), W):
            touched = set()
            for j, f, rb, cb in seq[r:r + W]:
                touched |= {(j, f, rb), (j, f, cb)}
            total += len(touched)
    return total


def test_headline_shape_rounds_share_input():
    L = plan([(512, 64)], 256, 1)
    assert L.wg_per_xcd == 32
    sched = schedule(L)
    for x, seq in sched.items():
        assert len(seq) == 8 * 36
        # rounds 0..7: one channel each; round 8: 4 leftover tiles of each channel
        for r in range(8):
            assert len({f for _, f, _, _ in seq[32 * r:32 * r + 32]}) == 1
        tail = seq[256:]
        for f in {f for _, f, _, _ in tail}:
            chunks = {c for _, g, rb, cb in tail if g == f for c in (rb, cb)}
            assert len(chunks) == 3
    # 8 chunks per channel is the floor; this order fetches 11 per channel
    assert chunk_streams(sched, 32) == 64 * 11


@pytest.mark.parametrize("nblk,tail,chunks", [(8, 4, 3), (8, 8, 4), (8, 12, 5), (8, 16, 6)])
def test_order_ends_compactly(nblk, tail, chunks):
    L = plan([(nblk * XG_TILE // 2, 1)], cus=8, wg_per_cu=1)
    L.wg_per_xcd = 1
    seq = schedule(L)[0]
    last = seq[-tail:]
    assert len({c for _, _, rb, cb in last for c in (rb, cb)}) == chunks
