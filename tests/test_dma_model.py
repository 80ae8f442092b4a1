"""CPU model of the addresses the int8 DMA kernel xg_xengine_rgd issues.

Restates, for every wave, lane, stage and DMA, the global buffer resource
(base, num_records), per-lane offset and LDS window (M0 + 16 lane) of
rgd_tile (xgpu_amd/csrc/xg_kernels.hip, the `dma` lambda), the transpose-read
offsets of its operands (rg_lane_off), and the epilogue's output stores
(rg_store_cells: address, bounds and triangle cut), and checks them against
the buffers they must stay inside:
  * every lane whose offset is in range (offset < num_records) reads 16 bytes
    inside the input block, and only bytes of the tile's own inputs in valid
    samples; every other lane is out of range (the hardware reads nothing);
  * every DMA writes a 1 KiB window inside the kernel's LDS allocation, and
    the DMAs of one stage tile that stage's buffer exactly once;
  * every operand read stays inside the stage buffer it reads;
  * every epilogue store lands inside the integration buffer, and each valid
    visibility is stored exactly once (the register-tile order's unused
    diagonal quadrant is zero-stored once, as the reference leaves it), for
    the 32x32x32 tiles' epilogue (rg_store_cells) and the 16x16x64 tiles'
    (rg16_store_cells, whose kernel shares rgd_tile's DMAs at 64-sample
    stages).
The shapes include the parity case kept off the GPU (96 stations, 2
channels, 257 samples, real/imag order) and the 512- and 2048-station shapes.
"""
import numpy as np
import pytest

TILE = 128          # inputs per tile side
ROW = 256           # LDS bytes per sample per tile side (int8)
NCHUNK = 16         # 16-byte chunks per sample row
OOR = 0x7fffff00    # offset given to lanes whose chunk is past the station count
CELL_GROUP, CELL_SCRATCH = 144, 16 * 144


def i32(v):
    """v as the kernel's 32-bit int arithmetic leaves it (two's complement wrap)."""
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


def kernel_nrec(nt, s0, spd, ts, row_bytes):
    """num_records of a DMA / buffer load as the kernels compute it, in 32-bit
    ints: rows clamped at 0 before the product (past the block nt - s0 is far
    below 0, and (nt - s0) * ts wrapped to a large positive range at
    tstride > ~7 MB before the clamp)."""
    rows = max(min(nt - s0, spd), 0)
    return i32(i32(i32(rows * ts) - ts) + row_bytes) if rows > 0 else 0


def dma_chunk(slot, srow):
    x = slot ^ (srow & 7)
    return x ^ ((x & 1) << 3)


def rg_slot(chunk, s7):
    return chunk ^ s7 ^ ((chunk & 1) << 3)


def lane_off(blk, lane):
    """rg_lane_off(2 blk, lane): per-lane LDS offset of a 32-real-row block."""
    g, kg, r = (lane >> 4) & 1, lane >> 5, lane & 15
    chunk = 2 * blk + g
    return (kg * 16 + (r >> 1)) * ROW + (rg_slot(chunk, r >> 1) << 4) + ((r & 1) << 3)


# element formats of the real-Gram kernels: LDS row bytes per sample and tile
# side, 16-byte chunks per row, inputs per chunk, bytes per input, and the
# chunk a DMA lane of sample row `srow` loads into its slot
FMT = {"i8": dict(row=256, nchunk=16, kin=8, bpi=2, chunk=lambda slot, srow: dma_chunk(slot, srow)),
       "f16": dict(row=512, nchunk=32, kin=4, bpi=4, chunk=lambda slot, srow: slot ^ ((srow & 3) << 2))}


def model_dma(ns, nf, nt, f, rb, cb, stg, nbuf, stages=None, fmt="i8"):
    """Checks the DMAs of tile (f, rb, cb) over `stages` (default: every stage
    the kernel issues, including those past the block)."""
    F = FMT[fmt]
    ROW, NCHUNK, kin, bpi = F["row"], F["nchunk"], F["kin"], F["bpi"]
    spd = 1024 // ROW                  # sample rows per wave DMA
    kdma = stg // (4 * spd)            # DMAs per wave per side per stage
    side_bytes, buf_bytes = stg * ROW, 2 * stg * ROW
    lds_bytes = nbuf * buf_bytes
    ts = nf * ns * 2 * bpi             # input bytes per sample
    total = nt * ts
    ninput = 2 * ns
    diag = rb == cb
    nstage = (nt + stg - 1) // stg
    row_bytes = min(ts, ROW)
    in_f = f * ns * 2 * bpi
    lanes = np.arange(64)
    if stages is None:
        stages = range(nstage + nbuf + 1)  # with an odd stage count, one more
    # LDS bytes written per stage by all waves: each stage buffer exactly once
    cover = {st: np.zeros(buf_bytes, dtype=np.int32) for st in stages}
    for wave in range(4):
        srow, slot = lanes // NCHUNK, lanes % NCHUNK
        chunk = F["chunk"](slot, wave * spd + srow)
        dvoff = srow * ts + 16 * chunk
        assert dvoff.max() < 2 ** 31  # an int in the kernel
        voff = {0: np.where(rb * TILE + kin * chunk < ninput, dvoff, OOR),
                1: np.where((not diag) & (cb * TILE + kin * chunk < ninput), dvoff, OOR)}
        for st in stages:
            for side in (0, 1):
                for j in range(kdma):
                    g = 4 * j + wave
                    s0 = st * stg + g * spd
                    rows = max(min(nt - s0, spd), 0)
                    nrec = kernel_nrec(nt, s0, spd, ts, row_bytes)
                    assert nrec == (rows * ts - ts + row_bytes if rows else 0) and nrec < OOR
                    base = in_f + (cb if side else rb) * ROW + s0 * ts
                    assert 0 <= base < (1 << 48)
                    m0 = (st & (nbuf - 1)) * buf_bytes + side * side_bytes + g * 1024
                    assert 0 <= m0 and m0 + 1024 <= lds_bytes
                    rel = m0 - (st & (nbuf - 1)) * buf_bytes
                    cover[st][rel:rel + 1024] += 1
                    v = voff[side]
                    live = v < nrec
                    if not live.any():
                        continue
                    lo = base + v[live]
                    assert lo.min() >= 0 and (lo + 16).max() <= total, (st, side, j)
                    # the bytes belong to this tile side's inputs of valid samples
                    t = lo // ts
                    e = lo - t * ts
                    blk = cb if side else rb
                    first = in_f + blk * ROW
                    assert (t < nt).all()
                    assert ((e >= first) & (e + 16 <= first + ROW)).all()
                    m = (e - in_f) // bpi  # first of the kin inputs read
                    assert (m + kin - 1 < ninput).all()
    for st in stages:
        assert (cover[st] == 1).all(), st


def check_lanes(lo, total, nt, ts, in_f, first_byte, span, ninput, bpi=2):
    """Lanes reading 16 bytes at byte offsets `lo` of the input block: inside
    the block, in valid samples, inside this tile side's `span` bytes of a
    sample row, and on valid inputs only (bpi bytes per input)."""
    assert lo.min() >= 0 and (lo + 16).max() <= total
    t = lo // ts
    e = lo - t * ts
    assert (t < nt).all()
    assert ((e >= first_byte) & (e + 16 <= first_byte + span)).all()
    m = (e - in_f) // bpi  # first of the 16 / bpi inputs read
    assert (m + 16 // bpi - 1 < ninput).all()


def model_rg_loads(ns, nf, nt, f, rb, cb, stg, fmt="i8"):
    """The register-staged kernels' loads (RgLoader: 256 / chunks-per-row
    sample rows per pass, STG / pass loads per thread and side) of tile
    (f, rb, cb) for every stage the tile and the next tile's prologue issue
    (up to nstage + 5)."""
    F = FMT[fmt]
    ROW, NCHUNK, kin, bpi = F["row"], F["nchunk"], F["kin"], F["bpi"]
    kpass = 256 // NCHUNK
    kloads = stg // kpass
    ts = nf * ns * 2 * bpi
    total, ninput = nt * ts, 2 * ns
    row_bytes = min(ts, ROW)
    in_f = f * ns * 2 * bpi
    nstage = (nt + stg - 1) // stg
    tid = np.arange(256)
    chunk, srow = tid % NCHUNK, tid // NCHUNK
    voff = srow * ts + 16 * chunk
    assert voff.max() < 2 ** 31
    v = {0: np.where(rb * TILE + kin * chunk < ninput, voff, OOR),
         1: np.where((rb != cb) & (cb * TILE + kin * chunk < ninput), voff, OOR)}
    for st in range(nstage + 7):
        for i in range(kloads):
            t0 = st * stg + kpass * i
            nrec = kernel_nrec(nt, t0, kpass, ts, row_bytes)
            for side in (0, 1):
                blk = cb if side else rb
                base = in_f + blk * ROW + t0 * ts
                live = v[side] < nrec
                if live.any():
                    check_lanes(base + v[side][live], total, nt, ts, in_f, in_f + blk * ROW, ROW,
                                ninput, bpi)


def kg_slot(chunk, s3):
    return chunk ^ (s3 << 2)


def model_kg_dma(ns, nf, nt, f, rb, cp):
    """The K-interleaved kernel's DMAs of pair tile (f, rb, cp): row side 4
    samples of 256 bytes per DMA, column side 2 samples of 512 bytes, 32-sample
    stages, 4 buffers; every stage issued (up to nstage + 4)."""
    ts = nf * ns * 4
    total, ninput = nt * ts, 2 * ns
    in_f = f * ns * 4
    arow, brow = min(ts, 256), min(ts, 512)
    nstage = (nt + 31) // 32
    lanes = np.arange(64)
    r, c = lanes >> 4, kg_slot(lanes & 15, lanes >> 4)
    voa = r * ts + 16 * c
    voa = np.where(rb * TILE + 8 * c < ninput, voa, OOR)
    vob = []
    for par in range(2):
        rr = lanes >> 5
        cc = kg_slot(lanes & 31, (2 * par + rr) & 3)
        vob.append(np.where(cp * 2 * TILE + 8 * cc < ninput, rr * ts + 16 * cc, OOR))
    for st in range(nstage + 5):
        for wave in range(4):
            for d in range(6):
                if d < 2:
                    g = 2 * wave + d
                    s0 = st * 32 + 4 * g
                    nrec = kernel_nrec(nt, s0, 4, ts, arow)
                    base, vv, first, span = in_f + rb * 256 + s0 * ts, voa, in_f + rb * 256, 256
                else:
                    j = 4 * wave + d - 2
                    s0 = st * 32 + 2 * j
                    nrec = kernel_nrec(nt, s0, 2, ts, brow)
                    base, vv, first, span = in_f + cp * 512 + s0 * ts, vob[(d - 2) & 1], in_f + cp * 512, 512
                live = vv < nrec
                if live.any():
                    check_lanes(base + vv[live], total, nt, ts, in_f, first, span, ninput)


@pytest.mark.parametrize("ns,nf,nt", [(96, 2, 257), (512, 8191, 64), (132, 2, 64), (256, 4, 1000),
                                      (8192, 8, 256)])
@pytest.mark.parametrize("stg", [64, 128])
def test_rg_loads_in_bounds(ns, nf, nt, stg):
    nblk = (2 * ns + TILE - 1) // TILE
    for f in sorted({0, nf - 1}):
        for rb in range(nblk):
            if nblk > 8 and rb not in (0, 1, nblk - 2, nblk - 1):
                continue
            for cb in range(rb + 1):
                model_rg_loads(ns, nf, nt, f, rb, cb, stg)


@pytest.mark.parametrize("ns,nf,nt", [(96, 2, 257), (512, 4095, 64), (130, 1, 64), (256, 4, 1000)])
def test_fp16_loads_in_bounds(ns, nf, nt):
    """The FP16 real-Gram kernels: register-staged rgh<32>, rgh<64> and the DMA
    kernels rgdh<16,8> and rgdh<32,4>, up to the largest FP16 block they take
    (NF=4095 at NSTATION=512: tstride 16.8 MB)."""
    assert 128 * nf * ns * 8 < 2 ** 31 - 4096
    nblk = (2 * ns + TILE - 1) // TILE
    for f in sorted({0, nf - 1}):
        for rb in range(nblk):
            for cb in range(rb + 1):
                for stg in (32, 64):
                    model_rg_loads(ns, nf, nt, f, rb, cb, stg, fmt="f16")
                model_dma(ns, nf, nt, f, rb, cb, 32, 4, fmt="f16")
                model_dma(ns, nf, nt, f, rb, cb, 16, 8, fmt="f16")


@pytest.mark.parametrize("ns,nf,nt", [(96, 2, 257), (512, 8191, 64), (132, 2, 64), (256, 4, 1000)])
def test_kg_dma_in_bounds(ns, nf, nt):
    nblk = (2 * ns + TILE - 1) // TILE
    for f in sorted({0, nf - 1}):
        for cp in range((nblk + 1) // 2):
            for rb in range(2 * cp, nblk):
                model_kg_dma(ns, nf, nt, f, rb, cp)


def model_reads(stg, nbuf, diag):
    side_bytes, buf_bytes = stg * ROW, 2 * stg * ROW
    for wave in range(4):
        wr, wc = wave >> 1, wave & 1
        for lane in range(64):
            for a in range(4):
                for ks in range(stg // 32):
                    o = ks * 32 * ROW + lane_off(4 * wr + a, lane)
                    assert 0 <= o and o + 8 * ROW + 8 <= side_bytes
                    ob = ks * 32 * ROW + lane_off(4 * wc + a, lane) + (0 if diag else side_bytes)
                    assert 0 <= ob and ob + 8 * ROW + 8 <= buf_bytes


def model_epilogue(ns, nf, f, rb, cb, order, accumulate, hits):
    """rg_store_cells for every storing wave and lane of tile (f, rb, cb);
    counts element stores in `hits` (flat output of 2 planes / interleaved)."""
    diag = rb == cb
    full = (not diag) and (rb * 64 + 64 <= ns)
    nb = ns * (ns + 1) // 2
    P = ns // 2
    RT = P * (P + 1) // 2
    if order == "regtile":
        plane = nf * RT * 16
    else:
        plane = nf * nb * 4
    for wave in range(4):
        wr, wc = wave >> 1, wave & 1
        if diag and wr < wc:
            continue  # the mirror wave stores nothing
        # epilogue LDS scratch of this wave
        assert (wr * 2 + wc) * CELL_SCRATCH + 15 * CELL_GROUP + 64 + 3 * 16 + 16 <= 64 * 1024
        for lane in range(64):
            hi, mj, bq = lane >> 5, (lane & 31) >> 2, lane & 3
            j = cb * 64 + 32 * wc + mj + 8 * bq
            for a in range(4):
                for q in range(4):
                    i = rb * 64 + 32 * wr + hi + 8 * a + 2 * q
                    if not full:
                        if i >= ns or j >= ns:
                            continue
                        if i < j and (order != "regtile" or (i >> 1) != (j >> 1) or accumulate):
                            continue
                    if order == "regtile":
                        base = 4 * ((f * 4 + 2 * (j & 1) + hi) * RT + (j >> 1))
                        R = i >> 1
                        p = base + 4 * (R * (R + 1) // 2)
                        p1 = p + plane
                        span = 4
                    elif order == "realimag":
                        p = 4 * (f * nb + j) + 4 * (i * (i + 1) // 2)
                        p1 = p + plane
                        span = 4
                    else:  # triangular, interleaved: 8 values per cell
                        p = 8 * (f * nb + j) + 8 * (i * (i + 1) // 2)
                        p1 = p + 4
                        span = 4
                    size = 2 * plane
                    assert 0 <= p and p + span <= size and 0 <= p1 and p1 + span <= size, (i, j)
                    hits[p:p + span] += 1
                    hits[p1:p1 + span] += 1


def model_epilogue16(ns, nf, f, rb, cb, order, accumulate, hits):
    """rg16_store_cells (the 16x16x64 kernel's epilogue) for every storing
    wave and lane of tile (f, rb, cb), counting element stores in `hits`."""
    diag = rb == cb
    full = (not diag) and (rb * 64 + 64 <= ns)
    nb = ns * (ns + 1) // 2
    P = ns // 2
    RT = P * (P + 1) // 2
    plane = nf * RT * 16 if order == "regtile" else nf * nb * 4
    for wave in range(4):
        wr, wc = wave >> 1, wave & 1
        if diag and wr < wc:
            continue
        for lane in range(64):
            bq = lane & 3
            assert (wr * 2 + wc) * CELL_SCRATCH + (lane >> 2) * CELL_GROUP + 128 <= 64 * 1024
            for bu in range(8):
                for bvg in range(2):
                    i = rb * 64 + 32 * wr + 4 * bu + (lane >> 4)
                    j = cb * 64 + 32 * wc + 16 * bvg + 4 * bq + ((lane & 15) >> 2)
                    if not full:
                        if i >= ns or j >= ns:
                            continue
                        if i < j and (order != "regtile" or (i >> 1) != (j >> 1) or accumulate):
                            continue
                    if order == "regtile":
                        R = i >> 1
                        p = 4 * ((f * 4 + 2 * (j & 1) + (i & 1)) * RT + R * (R + 1) // 2 + (j >> 1))
                        p1 = p + plane
                    else:
                        k = f * nb + i * (i + 1) // 2 + j
                        p = (8 if order == "tri" else 4) * k
                        p1 = p + 4 if order == "tri" else p + plane
                    size = 2 * plane
                    assert 0 <= p and p + 4 <= size and 0 <= p1 and p1 + 4 <= size, (i, j)
                    hits[p:p + 4] += 1
                    hits[p1:p1 + 4] += 1


SHAPES = [(96, 2, 0, "regtile"), (64, 2, 0, "tri"),  # empty blocks (a batch job of 0 samples)
          (96, 2, 257, "realimag"), (96, 2, 257, "regtile"), (96, 2, 257, "tri"),
          (64, 1, 100, "tri"), (132, 2, 64, "regtile"), (16, 2, 64, "regtile"),
          (256, 4, 1000, "regtile"), (6 * 4, 3, 65, "tri")]


@pytest.mark.parametrize("ns,nf,nt,order", SHAPES)
@pytest.mark.parametrize("stg,nbuf", [(32, 8), (64, 4)])
def test_dma_addresses_in_bounds(ns, nf, nt, order, stg, nbuf):
    nblk = (2 * ns + TILE - 1) // TILE
    for f in range(nf):
        for rb in range(nblk):
            for cb in range(rb + 1):
                model_dma(ns, nf, nt, f, rb, cb, stg, nbuf)


@pytest.mark.parametrize("ns,nf,nt", [(512, 64, 8192), (2048, 16, 4096), (8192, 8, 256),
                                      (32640, 1, 64)])
def test_dma_addresses_in_bounds_large(ns, nf, nt):
    """First and last channel, every tile, the stages at both ends of the
    block and past it (the address arithmetic is affine in the stage)."""
    stg, nbuf = 32, 8
    nstage = (nt + stg - 1) // stg
    stages = sorted(set(range(0, nbuf + 1)) | set(range(max(nstage - nbuf, 0), nstage + nbuf + 1)))
    nblk = (2 * ns + TILE - 1) // TILE
    for f in (0, nf - 1):
        for rb in range(nblk):
            for cb in range(rb + 1):
                if nblk > 8 and rb not in (0, 1, nblk - 2, nblk - 1):
                    continue
                model_dma(ns, nf, nt, f, rb, cb, stg, nbuf, stages)


def test_dma_addresses_at_the_offset_limit():
    """The largest block the engine gives the real-Gram kernels: 128 * tstride
    just below 2^31 - 4096 (NSTATION=512, NFREQUENCY=8191; tstride 16.8 MB).
    Every tile of the first and last channel, every stage."""
    ns, nf, nt = 512, 8191, 64
    assert 128 * nf * ns * 4 < 2 ** 31 - 4096 <= 128 * (nf + 1) * ns * 4
    nblk = (2 * ns + TILE - 1) // TILE
    for f in (0, nf - 1):
        for rb in range(nblk):
            for cb in range(rb + 1):
                for stg, nbuf in ((32, 8), (64, 4)):  # rgd<32,8>; rgd<64,4> and rgd16
                    model_dma(ns, nf, nt, f, rb, cb, stg, nbuf)


@pytest.mark.parametrize("ns,nf", [(512, 64), (512, 4096), (512, 8191), (2048, 2047), (96, 2)])
def test_buffer_ranges_never_wrap(ns, nf):
    """Every buffer load's num_records in every int8 kernel (DMA stages of
    rgd<32,8>, rgd<64,4> and rgd16, register-staged loads of rg<64> and
    rg<128>, the K-interleaved kernel's row and column DMAs) over every stage
    it issues, past the block included, for block lengths 1..300: the 32-bit
    computation equals the exact range and is 0 exactly when no sample of the
    load is valid.  At NF=8191 (tstride 16.8 MB) the unclamped form wrapped to
    ranges of ~2e9 bytes past the block."""
    ts = nf * ns * 4
    if 128 * ts >= 2 ** 31 - 4096:
        return  # the engine does not give such blocks to these kernels
    row_bytes = min(ts, ROW)
    forms = []  # (samples per stage, stages in flight past the block, samples per load)
    forms += [(32, 8, 4), (64, 4, 4)]         # rgd<32,8>, rgd<64,4> and rgd16: 4 rows per DMA
    forms += [(64, 2, 16), (128, 2, 16)]      # rg<64>, rg<128>: 16 rows per load pass
    forms += [(32, 4, 4), (32, 4, 2)]         # kg: row side 4, column side 2 rows per DMA
    for nt in list(range(1, 70)) + [255, 256, 257, 300]:
        for stg, ahead, spd in forms:
            nstage = (nt + stg - 1) // stg
            for st in range(nstage + ahead + 2):
                for s0 in range(st * stg, (st + 1) * stg, spd):
                    rows = max(min(nt - s0, spd), 0)
                    exact = rows * ts - ts + row_bytes if rows else 0
                    assert kernel_nrec(nt, s0, spd, ts, row_bytes) == exact
                    assert (exact == 0) == (rows == 0)


@pytest.mark.parametrize("stg,nbuf", [(32, 8), (64, 4)])
def test_operand_reads_inside_stage_buffer(stg, nbuf):
    for diag in (False, True):
        model_reads(stg, nbuf, diag)


@pytest.mark.parametrize("ns,nf,order,acc", [(96, 2, "realimag", 0), (96, 2, "regtile", 0),
                                             (96, 2, "tri", 1), (132, 2, "regtile", 0),
                                             (132, 2, "regtile", 1), (64, 1, "tri", 0),
                                             (6 * 4, 3, "tri", 0), (16, 2, "realimag", 1)])
@pytest.mark.parametrize("epilogue", [model_epilogue, model_epilogue16])
def test_epilogue_stores_in_bounds_and_exactly_once(ns, nf, order, acc, epilogue):
    nb = ns * (ns + 1) // 2
    P = ns // 2
    plane = nf * P * (P + 1) // 2 * 16 if order == "regtile" else nf * nb * 4
    hits = np.zeros(2 * plane, dtype=np.int64)
    nblk = (2 * ns + TILE - 1) // TILE
    for f in range(nf):
        for rb in range(nblk):
            for cb in range(rb + 1):
                epilogue(ns, nf, f, rb, cb, order, acc, hits)
    if order == "regtile" and acc:
        # accumulating calls skip quadrant "12" of diagonal register tiles,
        # which holds the zeros the first call stored (reference kernel.cuh:65-71)
        assert hits.max() == 1
        assert (hits == 1).sum() == 2 * nf * nb * 4
    else:
        assert (hits == 1).all(), (int((hits == 0).sum()), int((hits > 1).sum()))


def test_kernel_source_clamps_rows_before_the_range_product():
    """Every `rows * ts` range product in the kernel source multiplies a rows
    value clamped at 0 (the form kernel_nrec models), so no buffer range can
    wrap past the block."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "xgpu_amd", "csrc",
                            "xg_kernels.hip")).read()
    uses = [m.start() for m in re.finditer(r"rows \* ts", src)]
    assert len(uses) >= 5
    for u in uses:
        before = src[:u]
        d = before.rfind("const int rows = ")
        assert d >= 0
        line = before[d:before.find("\n", d)]
        assert re.match(r"const int rows = max\(min\([^;]*\), 0\);", line), line


def sweep_kernel(case, fmt, call, per):
    """The kernel the engine launches for call `call` of a sweep case
    (xg_launch_family, rg_stage_code): None for the complex-plane kernel,
    which takes nstation % 4 != 0 (int8), input that is not 16-byte aligned
    and tuning kernel 1."""
    ns, nf, nt, order, calls, t = case
    ts = nf * ns * 2 * FMT[fmt]["bpi"]
    if t["kernel"] == 1 or (fmt == "i8" and ns % 4) or (call * per * ts) % 16:
        return None
    if fmt == "i8" and t["kernel"] == 3:
        return ("kg",)
    if t["staging"] == 1:
        return ("rg", t["stage"])
    if fmt == "f16":  # rgdh<16,8> (stage 0 or 16) or rgdh<32,4>
        return ("dma", 32, 4, model_epilogue) if t["stage"] == 32 else ("dma", 16, 8, model_epilogue)
    if t["stage"] == 32:
        return ("dma", 32, 8, model_epilogue)
    return ("dma", 64, 4, model_epilogue16 if t["mfma"] == 2 else model_epilogue)


def _sweep_params():
    import shape_sweep as S
    return ([pytest.param(c, "i8", id="i8-" + S.case_id(c)) for c in S.I8_SWEEP] +
            [pytest.param(c, "f16", id="f16-" + S.case_id(c)) for c in S.F16_SWEEP])


@pytest.mark.parametrize("case,fmt", _sweep_params())
def test_shape_sweep_in_bounds(case, fmt):
    """Every case of the GPU parity sweep, modelled before it runs: per call,
    the loads of the kernel the engine picks (every tile of the first and last
    channel, every stage issued) and, for the real-Gram kernels, the output
    stores of every tile of every channel, each visibility exactly once."""
    import shape_sweep as S
    ns, nf, nt, order, calls, t = case
    lens = S.call_lengths(nt, calls)
    nblk = (2 * ns + TILE - 1) // TILE
    nb = ns * (ns + 1) // 2
    P = ns // 2
    plane = nf * P * (P + 1) // 2 * 16 if order == "regtile" else nf * nb * 4
    for call, n in enumerate(lens):
        k = sweep_kernel(case, fmt, call, lens[0])
        if k is None:
            continue
        for f in sorted({0, nf - 1}):
            if k[0] == "kg":  # pair tiles: column blocks 2 cp, 2 cp + 1; rows from 2 cp
                for cp in range((nblk + 1) // 2):
                    for rb in range(2 * cp, nblk):
                        model_kg_dma(ns, nf, n, f, rb, cp)
                continue
            for rb in range(nblk):
                for cb in range(rb + 1):
                    if k[0] == "rg":
                        model_rg_loads(ns, nf, n, f, rb, cb, k[1], fmt=fmt)
                    else:
                        model_dma(ns, nf, n, f, rb, cb, k[1], k[2], fmt=fmt)
        if k[0] == "kg":
            continue  # its epilogue (kg_store_cells) is covered by the GPU tests alone
        epilogue = k[3] if k[0] == "dma" else model_epilogue
        hits = np.zeros(2 * plane, dtype=np.int64)
        for f in range(nf):
            for rb in range(nblk):
                for cb in range(rb + 1):
                    epilogue(ns, nf, f, rb, cb, order, int(call > 0), hits)
        if order == "regtile" and call > 0:
            assert hits.max() == 1 and (hits == 1).sum() == 2 * nf * nb * 4
        else:
            assert (hits == 1).all()
