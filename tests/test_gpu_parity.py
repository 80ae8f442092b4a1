"""GPU parity: the HIP kernels against the CPU oracle, through the C-ABI.

Integer path: bit-exact against the exact int64 oracle.  FP32 plain path:
bit-identical to the reference's float CPU X-engine while sums stay below
2^24 (every partial sum is then an exactly representable integer), which
covers the cfg1 workload; the reference check metric is also applied with
1e-6.  FP16 path: against an FP64 oracle of the same fp16 inputs, relative
RMS error <= 1e-6 and max abs error <= 1e-5 * max|V| (fp32 accumulation of
exact fp16 products).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import workloads as W
from xgpu_amd import _lib
from conftest import flush
import shape_sweep as S

pytestmark = pytest.mark.gpu

ORDERS = {
    "regtile": X.REGISTER_TILE_TRIANGULAR_ORDER,
    "tri": X.TRIANGULAR_ORDER,
    "realimag": X.REAL_IMAG_TRIANGULAR_ORDER,
}


def to_tri(eng, raw, ns, nf, order):
    """Engine dump -> interleaved triangular (host reorder by the oracle)."""
    if order == X.TRIANGULAR_ORDER:
        return raw
    return oracle.reorder(order, raw, ns, nf)


def run_i8(ns, nf, nt, x, order=X.REGISTER_TILE_TRIANGULAR_ORDER, calls=1, **tuning):
    flush(f"[i8] ns={ns} nf={nf} nt={nt} order={order} calls={calls} {tuning}")
    eng = X.Engine(ns, nf, nt, compute=X.XGPU_INT8, order=order)
    if tuning:
        eng.set_tuning(**tuning)
    buf = X.DeviceBuffer.from_array(x)
    per = nt // calls
    step = per * nf * ns * 4
    for c in range(calls):
        n = per if c < calls - 1 else nt - per * (calls - 1)
        eng.correlate(buf.ptr + c * step, n)
    raw = eng.dump()
    wide = eng.dump(wide=True)
    eng.close()
    buf.free()
    return raw, wide


# (kernel, staging, stage): complex-plane kernel; real-Gram kernel with register
# staging (64 / 128 samples per stage) and with global->LDS DMA (32 / 64);
# K-interleaved kernel (pair tiles)
I8_VARIANTS = [(1, 0, 0), (2, 1, 64), (2, 1, 128), (2, 2, 32), (2, 2, 64), (3, 0, 0)]


@pytest.mark.parametrize("ns,nf,nt,order", [
    (16, 2, 64, "regtile"),
    (64, 3, 128, "regtile"),
    (64, 1, 100, "tri"),          # time tail inside a stage
    (96, 2, 257, "realimag"),     # tile tail (192 inputs), time tail
    (132, 2, 64, "regtile"),      # 264 inputs: 3 row blocks, partial last
    (2, 1, 1, "tri"),             # smallest
    (6, 2, 65, "tri"),            # nstation % 4 == 2
    (256, 4, 1000, "regtile"),
])
@pytest.mark.parametrize("kernel,staging,stage", I8_VARIANTS)
def test_i8_matches_exact_oracle(ns, nf, nt, order, kernel, staging, stage):
    """kernel 1: complex-plane int8 MFMA kernel; kernel 2: real-Gram kernel
    (used where nstation % 4 == 0, else the engine takes kernel 1) with
    register (1) or global->LDS DMA (2) staging; kernel 3: K-interleaved
    kernel (same inputs as kernel 2)."""
    if (kernel, staging, stage, ns, nf, nt, order) == (2, 2, 32, 96, 2, 257, "realimag"):
        pytest.skip("this case faulted the GPU twice with a since-reverted change to the DMA "
                    "kernel (operands made scalar with readfirstlane, per-order epilogue "
                    "specialization); kept off the GPU until the cause is shown")
    x = oracle.fill_uniform(nt * nf * ns * 4, ns * 1000 + nt, -128, 127)
    raw, wide = run_i8(ns, nf, nt, x, ORDERS[order], kernel=kernel, staging=staging, stage=stage)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(to_tri(None, raw, ns, nf, ORDERS[order]).astype(np.int64), ref)
    assert np.array_equal(to_tri(None, wide, ns, nf, ORDERS[order]), ref)


@pytest.mark.parametrize("ns,nf,nt,order", [
    (16, 2, 64, "regtile"), (64, 3, 128, "regtile"), (64, 1, 100, "tri"),
    (96, 2, 257, "tri"), (132, 2, 64, "regtile"), (256, 4, 1000, "realimag"),
    (512, 2, 8192, "regtile")])
def test_i8_mfma16_matches_exact_oracle(ns, nf, nt, order):
    """The DMA kernel with 16x16x64 MFMA tiles (tuning mfma 2, 64-sample
    stages), an autotuner candidate: bit-exact against the int64 oracle."""
    x = oracle.fill_uniform(nt * nf * ns * 4, ns * 1000 + nt + 7, -128, 127)
    raw, wide = run_i8(ns, nf, nt, x, ORDERS[order], kernel=2, staging=2, stage=64, mfma=2)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(to_tri(None, raw, ns, nf, ORDERS[order]).astype(np.int64), ref)
    assert np.array_equal(to_tri(None, wide, ns, nf, ORDERS[order]), ref)


@pytest.mark.parametrize("kernel,staging", [(1, 0), (2, 1), (2, 2), (3, 0)])
def test_i8_minus_128_everywhere(kernel, staging):
    ns, nf, nt = 64, 2, 4096
    x = np.full(nt * nf * ns * 4, -128, dtype=np.int8)
    raw, wide = run_i8(ns, nf, nt, x, X.TRIANGULAR_ORDER, kernel=kernel, staging=staging)
    v = raw.reshape(-1, 2).astype(np.int64)
    assert np.all(v[:, 0] == nt * 2 * 128 * 128)
    assert np.all(v[:, 1] == 0)
    # mixed extremes: x_i = -128 + 127i, x_j = 127 - 128i
    x = np.zeros((nt, nf, ns, 2, 2), dtype=np.int8)
    x[:, :, 0::2] = (-128, 127)
    x[:, :, 1::2] = (127, -128)
    raw, _ = run_i8(ns, nf, nt, x.reshape(-1), X.TRIANGULAR_ORDER, kernel=kernel, staging=staging)
    assert np.array_equal(raw.astype(np.int64), oracle.xengine_i64(x, ns, nf, nt))


def test_i8_int32_bound_block_exact():
    """One block of exactly 65535 samples of -128 + -128i: every accumulator
    reaches its int32 extreme (2 * 65535 * 128^2 = 2^31 - 2^15 per
    autocorrelation) without a spill, in both int8 kernels."""
    ns, nf, nt = 4, 1, 65535
    x = np.full(nt * nf * ns * 4, -128, dtype=np.int8)
    for kernel, staging in ((1, 0), (2, 1), (2, 2), (3, 0)):
        raw, wide = run_i8(ns, nf, nt, x, X.TRIANGULAR_ORDER, kernel=kernel, staging=staging)
        v = raw.reshape(-1, 2).astype(np.int64)
        assert np.all(v[:, 0] == nt * 2 * 128 * 128) and np.all(v[:, 1] == 0), (kernel, staging)
        assert np.array_equal(wide.reshape(-1, 2), v)


def test_i8_unaligned_input_falls_back_exact():
    """A block that is not 16-byte aligned goes to the complex-plane kernel."""
    ns, nf, nt = 64, 2, 200
    x = oracle.fill_uniform(nt * nf * ns * 4 + 8, 8, -128, 127)
    buf = X.DeviceBuffer.from_array(x)
    flush("[i8] input at +8 bytes")
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    eng.correlate(buf.ptr + 8, nt)
    ref = oracle.xengine_i64(x[8:], ns, nf, nt)
    assert np.array_equal(eng.dump().astype(np.int64), ref)
    eng.close()
    buf.free()


def test_i8_accumulates_across_calls_and_clear():
    ns, nf, nt = 64, 2, 640
    x = oracle.fill_uniform(nt * nf * ns * 4, 77, -128, 127)
    raw, _ = run_i8(ns, nf, nt, x, X.TRIANGULAR_ORDER, calls=5)
    assert np.array_equal(raw.astype(np.int64), oracle.xengine_i64(x, ns, nf, nt))
    flush("[i8] clear semantics")
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    assert not eng.dump().any()          # fresh engine dumps zeros
    eng.correlate(buf, nt)
    eng.clear()
    assert not eng.dump().any()          # clear is immediate
    eng.correlate(buf, nt)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(oracle.reorder(3000, eng.dump(), ns, nf).astype(np.int64), ref)
    eng.close()


def test_i8_long_integration_spills_to_int64():
    """Beyond 65535 samples the int32 buffer spills; the int64 dump is exact
    and the int32 dump equals the exact value modulo 2^32."""
    ns, nf, nt = 8, 1, 50000
    x = np.full(nt * nf * ns * 4, -128, dtype=np.int8)
    flush("[i8] spill ns=8 nt=3x50000")
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    for _ in range(3):
        eng.correlate(buf, nt)
    assert eng.samples == 3 * nt
    wide = eng.dump(wide=True).reshape(-1, 2)
    narrow = eng.dump().reshape(-1, 2)
    expect = 3 * nt * 2 * 128 * 128
    assert expect > 2**31
    assert np.all(wide[:, 0] == expect) and np.all(wide[:, 1] == 0)
    assert np.all(narrow[:, 0] == ((expect + 2**31) % 2**32) - 2**31)
    eng.close()
    # random data, one call longer than the bound
    ns, nf, nt = 4, 2, 70001
    x = oracle.fill_uniform(nt * nf * ns * 4, 5, -128, 127)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    eng.correlate(buf, nt)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(eng.dump(wide=True), ref)
    assert np.array_equal(eng.dump().astype(np.int64), ((ref + 2**31) % 2**32) - 2**31)
    eng.close()


def test_i8_device_reorder_dump():
    ns, nf, nt = 48, 3, 96
    x = oracle.fill_uniform(nt * nf * ns * 4, 3, -128, 127)
    for order in (X.REGISTER_TILE_TRIANGULAR_ORDER, X.REAL_IMAG_TRIANGULAR_ORDER):
        flush(f"[i8] device reorder order={order}")
        eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=order)
        buf = X.DeviceBuffer.from_array(x)
        eng.correlate(buf, nt)
        ref = oracle.xengine_i64(x, ns, nf, nt)
        assert np.array_equal(eng.dump(triangular=True).astype(np.int64), ref)
        assert np.array_equal(eng.dump(wide=True, triangular=True), ref)
        # to device memory: reordered straight into the destination, 16-byte
        # aligned (vector cell path) and 4 bytes off (scalar path)
        for wide, dt in ((False, np.int32), (True, np.int64)):
            n = 2 * eng.tri_len
            out = X.DeviceBuffer(n * 8 + 16)
            flags = X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR | (X.XGPU_DUMP_WIDE if wide else 0)
            for off in (0, 4):
                X.check(X.core().xgpuAmdDump(eng.handle, out.ptr + off, flags))
                got = out.download(dt, count=n, offset=off).astype(np.int64)
                assert np.array_equal(got, ref), (wide, off)
            out.free()
        eng.close()


@pytest.mark.parametrize("order", ["regtile", "realimag"])
def test_i8_spilled_wide_triangular_dump(order):
    """A spilled split-plane integration (register-tile or real/imag order)
    dumped in triangular order (spill + current part reordered in one pass),
    int64 and int32, to host and to device memory, aligned for paired stores
    and off by one element."""
    ns, nf, nt = 8, 2, 40000
    x = oracle.fill_uniform(nt * nf * ns * 4, 11, -128, 127)
    flush(f"[i8] spilled wide triangular dump ns=8 nf=2 3x40000 {order}")
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=ORDERS[order])
    buf = X.DeviceBuffer.from_array(x)
    for _ in range(3):
        eng.correlate(buf, nt)
    ref = 3 * oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(eng.dump(wide=True, triangular=True), ref)
    n = 2 * eng.tri_len
    out = X.DeviceBuffer(n * 8 + 16)
    flags = X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR | X.XGPU_DUMP_WIDE
    for off in (0, 8):
        X.check(X.core().xgpuAmdDump(eng.handle, out.ptr + off, flags))
        assert np.array_equal(out.download(np.int64, count=n, offset=off), ref), off
    narrow = ((ref + 2**31) % 2**32) - 2**31
    flags = X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR
    for off in (0, 4):
        X.check(X.core().xgpuAmdDump(eng.handle, out.ptr + off, flags))
        assert np.array_equal(out.download(np.int32, count=n, offset=off).astype(np.int64),
                              narrow), off
    out.free()
    assert np.array_equal(eng.dump(triangular=True).astype(np.int64), narrow)
    eng.close()


def test_i8_full_size_cfg2():
    """NS=512 x 2 pol, NF=64, NT=8192: exact on sampled baselines of every
    channel and on every visibility of 4 whole channels (int64 oracle), every
    autocorrelation against the input power, and every int8 kernel variant
    (16x16x64 MFMA tiles included) bit-identical to the default on every
    visibility."""
    ns, nf, nt = 512, 64, 8192
    flush("[i8] cfg2 full size")
    buf = X.DeviceBuffer(nt * nf * ns * 4)
    buf.fill_uniform_int8(2, -128, 127)
    x = buf.download(np.int8)
    assert np.array_equal(x[:1 << 20], oracle.fill_uniform(1 << 20, 2, -128, 127))
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    eng.correlate(buf, nt)
    tri = oracle.reorder(3000, eng.dump(), ns, nf).reshape(nf, -1, 4, 2).astype(np.int64)
    rng = np.random.default_rng(42)
    fij = [(f, i, j) for f in range(nf)
           for i, j in [(int(a), int(rng.integers(a + 1))) for a in rng.integers(0, ns, 6)]]
    fij += [(f, ns - 1, ns - 1) for f in range(0, nf, 7)] + [(f, ns - 1, 0) for f in range(0, nf, 5)]
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, i, j) in enumerate(fij):
        assert np.array_equal(tri[f, i * (i + 1) // 2 + j], ref[n]), (f, i, j)
    # every visibility of whole channels, the first and the last among them
    chans = [0, 21, 42, nf - 1]
    flush(f"[i8] cfg2 whole channels {chans}")
    whole = oracle.xengine_i64_channels(x, ns, nf, nt, chans)
    for c, f in enumerate(chans):
        assert np.array_equal(tri[f], whole[c]), f
    # every autocorrelation: sum_t |x|^2 per input, from the input alone
    xt = x.reshape(nt, nf * ns * 2 * 2)
    power = np.zeros(nf * ns * 2 * 2, dtype=np.int64)
    for t0 in range(0, nt, 512):
        c = xt[t0:t0 + 512].astype(np.int32)
        power += (c * c).sum(axis=0, dtype=np.int64)
    power = power.reshape(nf, ns, 2, 2).sum(axis=3)  # [f, s, p]
    autos = np.array([tri[:, i * (i + 1) // 2 + i, [0, 3], 0] for i in range(ns)])  # [s, f, p]
    assert np.array_equal(autos.transpose(1, 0, 2), power)
    flush("[i8] cfg2 full size, every int8 kernel variant against the default")
    first = eng.dump()
    for kernel, staging, stage, mfma in [v + (0,) for v in I8_VARIANTS] + [(2, 2, 64, 2)]:
        eng.set_tuning(kernel=kernel, staging=staging, stage=stage, mfma=mfma)
        eng.clear()
        eng.correlate(buf, nt)
        assert np.array_equal(eng.dump(), first), (kernel, staging, stage, mfma)
    eng.close()
    buf.free()


# ------------------------------------------------------------------ FP32 plain
def test_f32_matches_float_oracle_small():
    ns, nf, nt = 96, 3, 300
    x = oracle.fill_uniform(nt * nf * ns * 4, 12, -127, 127)
    flush("[f32] small")
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT32, order=X.TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    eng.correlate(buf, nt)
    got = eng.dump()
    assert np.array_equal(got, oracle.xengine_f32(x, ns, nf, nt))
    eng.close()


def test_f32_cfg1_through_xgpu_h():
    """cfg1 through the drop-in: xgpuInit -> xgpuCudaXengine(SYNCOP_DUMP) ->
    xgpuReorderMatrix, against the reference float CPU X-engine."""
    ns, nf, nt = 256, 10, 1024
    flush("[f32] cfg1 via libxgpu.so")
    lib = X.variant("libxgpu")
    ctx = _lib.XGPUContext()
    assert lib.xgpuInit(C.byref(ctx), 0) == 0
    info = _lib.XGPUInfo()
    lib.xgpuInfo(C.byref(info))
    assert (info.nstation, info.nfrequency, info.ntime) == (ns, nf, nt)
    x = oracle.fill_uniform(nt * nf * ns * 4, 1, -127, 127)
    C.memmove(ctx.array_h, x.ctypes.data, x.nbytes)
    assert lib.xgpuCudaXengine(C.byref(ctx), X.SYNCOP_DUMP) == 0
    mat = np.ctypeslib.as_array((C.c_float * (2 * info.matLength)).from_address(ctx.matrix_h))
    lib.xgpuReorderMatrix(ctx.matrix_h)
    got = mat[: 2 * info.triLength].copy()
    ref = oracle.xengine_f32(x, ns, nf, nt)
    nbad, maxerr = oracle.check(got, ref, ns, nf, tol=1e-6)
    assert nbad == 0, maxerr
    assert np.array_equal(got, ref)
    lib.xgpuFree(C.byref(ctx))
    assert ctx.internal is None and ctx.array_h is None


# ------------------------------------------------------------------ FP16 MFMA
# (kernel, staging, stage): complex-plane kernel; real-Gram kernel with
# register staging (32 / 64 samples per LDS stage) and with global->LDS DMA
# (16, the default, / 32)
F16_VARIANTS = [(1, 0, 0), (2, 1, 32), (2, 1, 64), (2, 2, 16), (2, 2, 32)]


def f16_input(nt, nf, ns, seed, sigma=16.0):
    rng = np.random.default_rng(seed)
    return rng.normal(0, sigma, (nt, nf, ns, 2, 2)).astype(np.float16)


def assert_f16_close(got, ref, ns, nf, nt, rms=True, calls=1):
    """fp32-accumulated visibilities (interleaved triangular) against the
    FP64 oracle of the same fp16 inputs, for nt samples correlated in `calls`
    calls (or host-path chunks) of S.call_lengths(nt, calls) samples.  Per
    baseline and polarization pair, |dV_ij| <= s 2^-24 sqrt(V_ii V_jj) with
    s = sum over calls of (ceil(n_c / 8) + 1), plus calls - 1: the fp16
    products are exact in fp32, every MFMA adds the products of at most 16
    samples (at least 8 in every MFMA but a call's last) to an fp32
    accumulator (one rounding of at most 2^-24 times the running sum of
    |products| per MFMA), the real-Gram combine rounds once more per call,
    each later call's result is added to the integration with one more, and
    sum_t |x_i x_j| <= sqrt(V_ii V_jj) (Cauchy-Schwarz).  Also max |dV| <=
    1e-5 max |V| and, for short blocks, relative RMS <= 1e-6."""
    import math
    steps = sum(math.ceil(n / 8) + 1 for n in S.call_lengths(nt, calls)) + calls - 1
    g = got.reshape(nf, -1, 4, 2).astype(np.float64)
    r = ref.reshape(nf, -1, 4, 2)
    dv = np.abs((g[..., 0] - r[..., 0]) + 1j * (g[..., 1] - r[..., 1]))
    mag = np.abs(r[..., 0] + 1j * r[..., 1])
    assert dv.max() <= 1e-5 * mag.max(), (dv.max(), mag.max())
    if rms:
        rel_rms = np.sqrt(np.mean(dv ** 2)) / np.sqrt(np.mean(mag ** 2))
        assert rel_rms <= 1e-6, rel_rms
    i = np.repeat(np.arange(ns), np.arange(1, ns + 1))
    j = np.concatenate([np.arange(k + 1) for k in range(ns)])
    di = i * (i + 1) // 2 + i
    dj = j * (j + 1) // 2 + j
    auto = r[:, :, [0, 3], 0]  # [f, baseline, (p=0, p=1)] real autocorrelations
    for p1 in (0, 1):
        for p2 in (0, 1):
            norm = np.sqrt(auto[:, di, p1] * auto[:, dj, p2])
            worst = (dv[:, :, 2 * p1 + p2] / norm).max()
            assert worst <= steps * 2.0 ** -24, (p1, p2, worst, steps)


def run_f16(ns, nf, nt, x16, order, calls=1, **tuning):
    flush(f"[f16] ns={ns} nf={nf} nt={nt} order={order} calls={calls} {tuning}")
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=order)
    if tuning:
        eng.set_tuning(**tuning)
    buf = X.DeviceBuffer.from_array(x16)
    per = nt // calls
    step = per * nf * ns * 8
    for c in range(calls):
        n = per if c < calls - 1 else nt - per * (calls - 1)
        eng.correlate(buf.ptr + c * step, n)
    raw = eng.dump()
    wide = eng.dump(wide=True)
    eng.close()
    buf.free()
    return to_tri(None, raw, ns, nf, order), to_tri(None, wide, ns, nf, order)


@pytest.mark.parametrize("ns,nf,nt,order,calls", [
    (64, 2, 512, "tri", 1),
    (100, 3, 300, "regtile", 1),    # ragged tile edge, partial last stage
    (256, 1, 97, "realimag", 2),    # two tiles per side, accumulate
    (2, 1, 33, "tri", 1),           # one valid chunk
    (132, 2, 200, "regtile", 2),    # 3 tiles per side, accumulate, ragged edge
    (130, 1, 64, "realimag", 1),    # nstation % 4 == 2
    (92, 3, 7, "realimag", 2),      # 3 + 4 samples: one MFMA step per call, accumulated
])
@pytest.mark.parametrize("kernel,staging,stage", F16_VARIANTS)
def test_f16_against_f64_oracle(ns, nf, nt, order, calls, kernel, staging, stage):
    x16 = f16_input(nt, nf, ns, 4 + ns + nt)
    got, wide = run_f16(ns, nf, nt, x16, ORDERS[order], calls, kernel=kernel, staging=staging,
                        stage=stage)
    ref = oracle.xengine_f64_f16(x16, ns, nf, nt)
    assert_f16_close(got, ref, ns, nf, nt, calls=calls)
    assert np.array_equal(wide, got.astype(np.float64))


def test_f16_variants_agree_cfg2_shape():
    """NSTATION=512 x 2 pol, NFREQUENCY=64 over 1024 samples: the real-Gram
    kernels (both stage depths, both stagings) agree with each other bit for
    bit (same MFMA sequence per accumulator) and with the complex-plane kernel
    within the fp32 accumulation bound."""
    ns, nf, nt = 512, 64, 1024
    x16 = f16_input(nt, nf, ns, 11)
    outs = {}
    for kernel, staging, stage in F16_VARIANTS:
        outs[(kernel, staging, stage)], _ = run_f16(ns, nf, nt, x16, X.TRIANGULAR_ORDER, 1,
                                                    kernel=kernel, staging=staging, stage=stage)
    assert np.array_equal(outs[(2, 1, 32)], outs[(2, 1, 64)])
    assert np.array_equal(outs[(2, 1, 32)], outs[(2, 2, 32)])
    assert np.array_equal(outs[(2, 1, 32)], outs[(2, 2, 16)])
    chans = [0, 63]
    ref = oracle.xengine_f64_f16(np.ascontiguousarray(x16[:, chans]), ns, len(chans), nt)
    for v, got in outs.items():
        g = got.reshape(nf, -1)[chans].reshape(-1)
        assert_f16_close(g, ref, ns, len(chans), nt)


def test_f16_spill_to_f64():
    """70000 samples in one call: the fp32 integration spills into the FP64
    one past 65536 samples; the wide dump matches the FP64 oracle within the
    accumulation bound."""
    ns, nf, nt = 8, 1, 70000
    x16 = f16_input(nt, nf, ns, 21)
    flush("[f16] spill")
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x16)
    eng.correlate(buf, nt)
    wide = eng.dump(wide=True)
    assert eng.samples == nt
    eng.close()
    buf.free()
    ref = oracle.xengine_f64_f16(x16, ns, nf, nt)
    assert_f16_close(wide, ref, ns, nf, nt, rms=False)


@pytest.mark.parametrize("order", ["regtile", "realimag"])
def test_f16_spilled_triangular_dump(order):
    """A spilled fp16 integration (3 x 40000 samples: the fp32 part spills
    into the FP64 one past 65536 samples) of a split-plane order dumped in
    triangular order by the one-pass combine + reorder: double (wide) and
    float, to host and to device memory, aligned and off by one element.
    Against the FP64 oracle: normalized error |dV_ij| / sqrt(V_ii V_jj) <=
    (3 NT / 16 + 4) 2^-24 (fp32 accumulation of each spill window, the
    combine, and for the float dump its final rounding)."""
    ns, nf, nt = 8, 2, 40000
    x16 = f16_input(nt, nf, ns, 22, sigma=4.0)
    flush(f"[f16] spilled triangular dump {order}")
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=ORDERS[order])
    buf = X.DeviceBuffer.from_array(x16)
    for _ in range(3):
        eng.correlate(buf, nt)
    ref = 3 * oracle.xengine_f64_f16(x16, ns, nf, nt)
    bound = (3 * nt / 16 + 4) * 2.0 ** -24
    outs = [eng.dump(wide=True, triangular=True), eng.dump(triangular=True)]
    n = 2 * eng.tri_len
    out = X.DeviceBuffer(n * 8 + 16)
    for wide, dt, offs in ((True, np.float64, (0, 8)), (False, np.float32, (0, 4))):
        flags = X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR | (X.XGPU_DUMP_WIDE if wide else 0)
        for off in offs:
            X.check(X.core().xgpuAmdDump(eng.handle, out.ptr + off, flags))
            outs.append(out.download(dt, count=n, offset=off))
    out.free()
    eng.close()
    buf.free()
    for k, got in enumerate(outs):
        st = W.vis_error_stats(got, ref, ns)
        assert st["max_norm_err"] <= bound, (k, st)
    # host and device dumps of one kind are the same values
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[0], outs[3])
    assert np.array_equal(outs[1], outs[4]) and np.array_equal(outs[1], outs[5])


@pytest.mark.parametrize("compute", ["i8", "f16"])
def test_triangular_device_dumps_every_state(compute):
    """A triangular-order engine dumps into device memory through the kernel
    that forms the dump, writing straight into the destination: a zero fill
    after a clear, the widening before any spill, the spill + current
    combine after one (3 x 40000 samples), wide and narrow.  Destinations
    16-byte aligned (4 elements per thread) and off by one element (one per
    thread), filled with garbage first.  int8 against k x the int64 oracle
    exactly (narrow: modulo 2^32); fp16 device dumps equal to the host dumps
    of the same state, which are within (k NT / 16 + 4) 2^-24 of the FP64
    oracle, normalized."""
    ns, nf, nt = 8, 2, 40000
    i8 = compute == "i8"
    if i8:
        x = oracle.fill_uniform(nt * nf * ns * 4, 31, -128, 127)
        ref1 = oracle.xengine_i64(x, ns, nf, nt)
    else:
        x = f16_input(nt, nf, ns, 32, sigma=4.0)
        ref1 = oracle.xengine_f64_f16(x, ns, nf, nt)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8 if i8 else X.XGPU_FLOAT16,
                   order=X.TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    n = ref1.size
    out = X.DeviceBuffer(n * 8 + 16)
    wdt, ndt = (np.int64, np.int32) if i8 else (np.float64, np.float32)
    done = 0
    for k in (0, 1, 3):
        flush(f"[{compute}] triangular device dumps after {k} x {nt} samples")
        if k == 0:
            eng.clear()
        while done < k:
            eng.correlate(buf, nt)
            done += 1
        host = {True: eng.dump(wide=True), False: eng.dump()}
        for wide, dt, offs in ((True, wdt, (0, 8)), (False, ndt, (0, 4))):
            flags = X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR | (X.XGPU_DUMP_WIDE if wide else 0)
            for off in offs:
                out.fill_uniform_int8(40 + off, -127, 127)
                X.check(X.core().xgpuAmdDump(eng.handle, out.ptr + off, flags))
                got = out.download(dt, count=n, offset=off)
                assert np.array_equal(got, host[wide]), (k, wide, off)
        if i8:
            exact = k * ref1
            assert np.array_equal(host[True], exact), k
            assert np.array_equal(host[False].astype(np.int64),
                                  ((exact + 2**31) % 2**32) - 2**31), k
        elif k == 0:
            assert not host[True].any() and not host[False].any()
        else:
            bound = (k * nt / 16 + 4) * 2.0 ** -24
            for wide in (True, False):
                st = W.vis_error_stats(host[wide], k * ref1, ns)
                assert st["max_norm_err"] <= bound, (k, wide, st)
    out.free()
    buf.free()
    eng.close()


# ------------------------------------------------------------------ hipGraph steps
def test_device_path_graph_replays():
    """graph knob on: a device-path call (layout conversion + launch) is
    captured once per launch and replayed.  Several replays, with eager work
    and synchronizations between them, each checked against the oracle:
    store after a clear, accumulate, new data in the same buffer, a
    time-swizzled engine, and a call past the int32 bound (issued eagerly with
    a spill)."""
    ns, nf, nt = 64, 3, 4096
    xs = [oracle.fill_uniform(nt * nf * ns * 4, 700 + k, -128, 127) for k in range(3)]
    refs = [oracle.xengine_i64(x, ns, nf, nt) for x in xs]
    buf = X.DeviceBuffer(xs[0].nbytes)
    other = X.DeviceBuffer.from_array(xs[2])
    plain = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)  # eager work
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    eng.set_tuning(graph=1)
    flush("[graph] device path replays")
    for rep in range(4):
        k = rep % 2
        buf.upload(xs[k])
        eng.clear()
        eng.correlate(buf, nt)                      # store step
        plain.clear()
        plain.correlate(other, nt)                  # eager work between replays
        assert np.array_equal(eng.dump(wide=True, triangular=True), refs[k]), rep
        eng.correlate(buf, nt)                      # accumulate step
        X.core().xgpuAmdDeviceSynchronize()
        assert np.array_equal(eng.dump(wide=True, triangular=True), 2 * refs[k]), rep
        assert np.array_equal(plain.dump().astype(np.int64), refs[2]), rep
    # past the int32 bound: 16 calls of 4096 = 65536 samples, the last spills
    eng.clear()
    buf.upload(xs[0])
    for _ in range(16):
        eng.correlate(buf, nt)
    assert np.array_equal(eng.dump(wide=True, triangular=True), 16 * refs[0])
    # a time-swizzled engine: the unswizzle is part of the captured step
    sw = X.DeviceBuffer.from_array(oracle.swizzle(xs[1], ns, nf, nt))
    es = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER, time_swizzle=4)
    es.set_tuning(graph=1)
    for rep in range(3):
        es.clear()
        es.correlate(sw, nt)
        assert np.array_equal(es.dump(wide=True), refs[1]), rep
    for b in (buf, other, sw):
        b.free()
    for e in (plain, eng, es):
        e.close()


def test_batch_graph_replays():
    """graph knob on the batch's first engine: one ragged launch captured and
    replayed; between replays the inputs are refilled, other engines work
    eagerly, and each replay's integrations are checked against the oracle."""
    shapes = [(48, 3, 512), (96, 2, 1000), (20, 4, 300)]
    engines, bufs = [], []
    for ns, nf, nt in shapes:
        engines.append(X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER))
        bufs.append(X.DeviceBuffer(nt * nf * ns * 4))
    engines[0].set_tuning(graph=1)
    items = [(e, b, s[2]) for e, b, s in zip(engines, bufs, shapes)]
    extra = X.Engine(64, 2, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    xe = oracle.fill_uniform(256 * 2 * 64 * 4, 799, -128, 127)
    be = X.DeviceBuffer.from_array(xe)
    flush("[graph] batch replays")
    for rep in range(4):
        xs = [oracle.fill_uniform(nt * nf * ns * 4, 800 + 10 * rep + k, -128, 127)
              for k, (ns, nf, nt) in enumerate(shapes)]
        for b, x in zip(bufs, xs):
            b.upload(x)
        for e in engines:
            e.clear()
        X.correlate_batch(items)
        extra.clear()
        extra.correlate(be, 256)
        X.correlate_batch(items)                 # accumulate replay
        for k, (e, (ns, nf, nt)) in enumerate(zip(engines, shapes)):
            assert np.array_equal(e.dump(wide=True), 2 * oracle.xengine_i64(xs[k], ns, nf, nt)), (rep, k)
        assert np.array_equal(extra.dump().astype(np.int64), oracle.xengine_i64(xe, 64, 2, 256))
    for b in bufs + [be]:
        b.free()
    for e in engines + [extra]:
        e.close()


# ------------------------------------------------------------------ batching
@pytest.mark.parametrize("kernel", [0, 3])
def test_batch_ragged_equals_alone(kernel):
    """One persistent launch for 4 engines of different shapes (square tiles
    with the default kernel, pair tiles with kernel 3): each equals its
    oracle, and a second batch accumulates."""
    shapes = [(64, 3, 256), (96, 2, 512), (200, 1, 300), (16, 5, 64)]
    engines, bufs, refs = [], [], []
    for k, (ns, nf, nt) in enumerate(shapes):
        x = oracle.fill_uniform(nt * nf * ns * 4, 50 + k, -128, 127)
        engines.append(X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER))
        if kernel:
            engines[-1].set_tuning(kernel=kernel)
        bufs.append(X.DeviceBuffer.from_array(x))
        refs.append(oracle.xengine_i64(x, ns, nf, nt))
    flush(f"[i8] ragged batch of 4, kernel {kernel}")
    X.correlate_batch([(e, b, s[2]) for e, b, s in zip(engines, bufs, shapes)])
    for e, r in zip(engines, refs):
        assert np.array_equal(e.dump().astype(np.int64), r)
    # second batch accumulates
    X.correlate_batch([(e, b, s[2]) for e, b, s in zip(engines, bufs, shapes)])
    for e, r in zip(engines, refs):
        assert np.array_equal(e.dump(wide=True), 2 * r)
        e.close()


def test_batch_of_the_maximum_16_jobs():
    """16 engines of different station counts, channel counts, block lengths
    and output orders in one launch (the most a batch takes): each equals its
    oracle; a second batch of every other engine accumulates."""
    rng = np.random.default_rng(16)
    orders = [X.TRIANGULAR_ORDER, X.REAL_IMAG_TRIANGULAR_ORDER, X.REGISTER_TILE_TRIANGULAR_ORDER]
    cases = []
    for k in range(16):
        ns = int(rng.choice([4, 8, 16, 28, 64, 96, 132, 200]))
        nf, nt = int(rng.integers(1, 6)), int(rng.integers(1, 400))
        order = orders[k % 3]
        x = oracle.fill_uniform(nt * nf * ns * 4, 700 + k, -128, 127)
        eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=order)
        cases.append((eng, X.DeviceBuffer.from_array(x), nt, oracle.xengine_i64(x, ns, nf, nt),
                      ns, nf, order))
    flush(f"[i8] batch of 16: {[(c[4], c[5], c[2]) for c in cases]}")
    X.correlate_batch([(e, b, nt) for e, b, nt, *_ in cases])
    for e, b, nt, ref, ns, nf, order in cases:
        assert np.array_equal(to_tri(None, e.dump(wide=True), ns, nf, order), ref), (ns, nf, nt)
    X.correlate_batch([(e, b, nt) for e, b, nt, *_ in cases[::2]])
    for k, (e, b, nt, ref, ns, nf, order) in enumerate(cases):
        assert np.array_equal(to_tri(None, e.dump(wide=True), ns, nf, order), (2 - k % 2) * ref)
        e.close()
        b.free()


# ------------------------------------------------------------------ host path + graphs
def _host_case(ns=64, nf=4, nt=512, ntp=64, seed=9):
    xs = [oracle.fill_uniform(nt * nf * ns * 4, seed + k, -128, 127) for k in range(3)]
    refs = [oracle.xengine_i64(x, ns, nf, nt) for x in xs]
    return ns, nf, nt, ntp, xs, refs


@pytest.mark.parametrize("graph", [2, 1])
def test_host_path_pipeline_and_graph_replays(graph):
    """xgpuAmdCorrelateHost (the xgpuCudaXengine pipeline): chunked transfers and
    kernels; with graphs on, each call is captured once and replayed.  Every
    replay is checked against the oracle, with eager work and syncs between."""
    ns, nf, nt, ntp, xs, refs = _host_case()
    flush(f"[host] graph={graph}")
    eng = X.Engine(ns, nf, nt, ntp, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    eng.set_tuning(graph=graph)
    hb = [X.HostBuffer(x.nbytes) for x in xs]
    for h, x in zip(hb, xs):
        h.array[:] = x
    out = np.zeros(2 * eng.mat_len, dtype=np.int32)
    lib = X.core()
    for rep in range(3):
        eng.clear()
        for k in range(3):  # three calls accumulate
            sync = X.SYNCOP_DUMP if k == 2 else X.SYNCOP_SYNC_TRANSFER
            X.check(lib.xgpuAmdCorrelateHost(eng.handle, hb[k].ptr, sync, out.ctypes.data))
        assert np.array_equal(out.astype(np.int64), refs[0] + refs[1] + refs[2]), rep
        # eager device-path work on the same engine between replays
        eng.clear()
        buf = X.DeviceBuffer.from_array(xs[rep])
        eng.correlate(buf, nt)
        assert np.array_equal(eng.dump().astype(np.int64), refs[rep])
        buf.free()
    assert eng.get_tuning()["graph"] == graph
    eng.close()
    for h in hb:
        h.free()


@pytest.mark.parametrize("graph", [2, 1])
def test_f16_host_path_batch_and_graph_steps(graph):
    """The FP16 path through the host-buffer pipeline (chunked transfers and
    kernels, captured in hipGraphs with the graph knob), the device path with
    the same knob, and a ragged batch of three FP16 engines: each against the
    FP64 oracle within the fp32 accumulation bound."""
    ns, nf, nt, ntp = 64, 3, 512, 128
    x = f16_input(nt, nf, ns, 81 + graph)
    ref = oracle.xengine_f64_f16(x, ns, nf, nt)
    flush(f"[f16] host path, graph={graph}")
    eng = X.Engine(ns, nf, nt, ntp, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
    eng.set_tuning(graph=graph)
    hb = X.HostBuffer(x.nbytes)
    hb.array[:] = x.reshape(-1).view(np.int8)
    out = np.zeros(2 * eng.mat_len, dtype=np.float32)
    lib = X.core()
    for rep in range(3):
        eng.clear()
        X.check(lib.xgpuAmdCorrelateHost(eng.handle, hb.ptr, X.SYNCOP_DUMP, out.ctypes.data))
        assert_f16_close(out, ref, ns, nf, nt, calls=nt // ntp)
    buf = X.DeviceBuffer.from_array(x)
    for rep in range(3):  # device path: captured once with the knob on, then replayed
        eng.clear()
        eng.correlate(buf, nt)
        assert_f16_close(eng.dump(), ref, ns, nf, nt)
    eng.close()
    hb.free()
    # ragged FP16 batch
    shapes = [(64, 3, 512), (96, 2, 100), (130, 1, 64)]
    cases = []
    for k, (bns, bnf, bnt) in enumerate(shapes):
        xk = f16_input(bnt, bnf, bns, 90 + k)
        e = X.Engine(bns, bnf, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
        e.set_tuning(graph=graph)
        cases.append((e, X.DeviceBuffer.from_array(xk), bnt, oracle.xengine_f64_f16(xk, bns, bnf, bnt),
                      bns, bnf))
    flush(f"[f16] ragged batch, graph={graph}")
    for rep in range(2):
        for e, *_ in cases:
            e.clear()
        X.correlate_batch([(e, b, t) for e, b, t, *_ in cases])
        for e, b, t, r, bns, bnf in cases:
            assert_f16_close(e.dump(), r, bns, bnf, t)
    for e, b, *_ in cases:
        e.close()
        b.free()
    buf.free()


def test_autotune_keeps_results_exact(tmp_path):
    """Every candidate the autotuner times is exact; the tuner leaves the
    caller's integration as it was (candidates run on a scratch engine); the
    winner goes to the tuning cache and an engine created later loads it."""
    ns, nf, nt = 128, 8, 1024
    x = oracle.fill_uniform(nt * nf * ns * 4, 31, -128, 127)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    buf = X.DeviceBuffer.from_array(x)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    cache = str(tmp_path / "tuning.txt")
    flush("[autotune] ns=128 nf=8 nt=1024")
    eng.correlate(buf, nt)                       # an integration in progress
    best, ms = eng.autotune(buf, nt, warmup=1, reps=2, cache_path=cache)
    assert ms > 0 and best["wg_per_cu"] in (1, 2) and best["prefetch"] in (1, 2)
    assert np.array_equal(eng.dump(wide=True), ref)          # untouched by the tuner
    eng.correlate(buf, nt)
    assert np.array_equal(eng.dump(wide=True), 2 * ref)      # and it continues
    assert best["kernel"] in (1, 2, 3) and best["stage"] in (0, 32, 64, 128)
    assert best["mfma"] in (1, 2)
    for wg in (1, 2):           # every candidate is exact
        for pf in (1, 2):
            eng.set_tuning(wg, pf, kernel=1)
            eng.clear()
            eng.correlate(buf, nt)
            assert np.array_equal(eng.dump().astype(np.int64), ref), (wg, pf)
    for kernel, staging, stage in I8_VARIANTS[1:]:
        eng.set_tuning(kernel=kernel, staging=staging, stage=stage)
        eng.clear()
        eng.correlate(buf, nt)
        assert np.array_equal(eng.dump().astype(np.int64), ref), (kernel, staging, stage)
    eng.set_tuning(kernel=2, staging=2, stage=64, mfma=2)
    eng.clear()
    eng.correlate(buf, nt)
    assert np.array_equal(eng.dump().astype(np.int64), ref), "mfma 16x16x64"
    eng2 = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    assert eng2.load_tuning(cache, nt)
    got = eng2.get_tuning()
    for k in ("wg_per_cu", "prefetch", "kernel", "mfma"):
        assert got[k] == best[k], k
    assert got["stage"] == best["stage"] or best["kernel"] == 1
    assert got["staging"] == best["staging"] or best["kernel"] == 1
    assert not eng2.load_tuning(cache, nt + 1)
    eng.close()
    eng2.close()
    buf.free()


def test_autotune_fp16_candidates():
    """The FP16 autotuner's candidates (complex-plane kernel; real-Gram kernel
    with register staging at 32 / 64 and DMA staging at 16 / 32 samples per
    stage) all run; the winner is kept and its result is within the fp32
    accumulation bound of the FP64 oracle."""
    ns, nf, nt = 128, 4, 512
    x16 = f16_input(nt, nf, ns, 23)
    ref = oracle.xengine_f64_f16(x16, ns, nf, nt)
    buf = X.DeviceBuffer.from_array(x16)
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
    flush("[autotune] fp16 ns=128 nf=4 nt=512")
    assert eng.get_tuning()["staging"] == 2 and eng.get_tuning()["stage"] == 16
    best, ms = eng.autotune(buf, nt, warmup=1, reps=2)
    assert ms > 0 and best["kernel"] in (1, 2)
    if best["kernel"] == 2:
        assert (best["staging"], best["stage"]) in ((1, 32), (1, 64), (2, 16), (2, 32))
    eng.clear()
    eng.correlate(buf, nt)
    assert_f16_close(eng.dump(), ref, ns, nf, nt)
    eng.close()
    buf.free()


def test_autotune_host_path_time_pipe(tmp_path):
    """Host-path tuner over the time-pipe length x graph replay: the winner
    divides NTIME, every setting gives the same exact integration, and the
    cache carries it to a new engine."""
    ns, nf, nt, ntp = 64, 4, 1024, 128
    x = oracle.fill_uniform(nt * nf * ns * 4, 41, -128, 127)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    cache = str(tmp_path / "tuning.txt")
    eng = X.Engine(ns, nf, nt, ntp, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    flush("[autotune] host path")
    got = eng.correlate_host(x, X.SYNCOP_DUMP)   # an integration in progress
    assert np.array_equal(got.astype(np.int64), ref)
    best, ms = eng.autotune_host(x, warmup=1, reps=2, cache_path=cache)
    assert ms > 0 and nt % best["ntimepipe"] == 0 and best["graph"] in (1, 2)
    assert np.array_equal(eng.dump(wide=True), ref)          # untouched by the tuner
    for p in (nt, nt // 2, nt // 8, 32):
        for graph in (1, 2):
            eng.set_tuning(ntimepipe=p, graph=graph)
            eng.clear()
            got = eng.correlate_host(x, X.SYNCOP_DUMP)
            assert np.array_equal(got.astype(np.int64), ref), (p, graph)
    # a device-path line is needed for the key the loader looks up first
    buf = X.DeviceBuffer.from_array(x)
    eng.autotune(buf, nt, warmup=1, reps=1, cache_path=cache)
    eng2 = X.Engine(ns, nf, nt, ntp, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    assert eng2.load_tuning(cache, nt)
    assert eng2.get_tuning()["ntimepipe"] == best["ntimepipe"]
    assert eng2.get_tuning()["graph"] == best["graph"]
    eng.close()
    eng2.close()
    buf.free()


# ------------------------------------------------------------------ full-size shapes
def test_i8_cfg3_shape_exact_by_properties():
    """NSTATION=2048 x 2 pol, NFREQUENCY=16, NTIME=4096 (1 GiB of int32 output),
    input round(N(0, 20^2)) plus a tone of amplitude 8 with a random phase per
    station (the cfg3 generator), clipped to [-127, 127].  Exact, via the
    device reorder to triangular order, on sampled baselines of every channel
    and every visibility of 4 whole channels (int64 oracle), and on every
    autocorrelation (input power)."""
    ns, nf, nt = 2048, 16, 4096
    flush("[i8] cfg3 shape")
    x = W.cfg3_input(ns, nf, nt, seed=3)
    rng = np.random.default_rng(33)
    x = x.reshape(-1)
    buf = X.DeviceBuffer.from_array(x)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    eng.correlate(buf, nt)
    tri = eng.dump(triangular=True).reshape(nf, -1, 4, 2).astype(np.int64)
    eng.close()
    buf.free()
    fij = [(f, i, j) for f in range(nf)
           for i, j in [(int(a), int(rng.integers(a + 1))) for a in rng.integers(0, ns, 8)]]
    fij += [(f, ns - 1, 0) for f in range(nf)] + [(f, ns - 1, ns - 2) for f in range(nf)]
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, i, j) in enumerate(fij):
        assert np.array_equal(tri[f, i * (i + 1) // 2 + j], ref[n]), (f, i, j)
    # every visibility of whole channels, the first and the last among them
    chans = [0, 5, 10, nf - 1]
    flush(f"[i8] cfg3 whole channels {chans}")
    whole = oracle.xengine_i64_channels(x, ns, nf, nt, chans)
    for c, f in enumerate(chans):
        assert np.array_equal(tri[f], whole[c]), f
    xt = x.reshape(nt, nf * ns * 4)
    power = np.zeros(nf * ns * 4, dtype=np.int64)
    for t0 in range(0, nt, 256):
        c = xt[t0:t0 + 256].astype(np.int32)
        power += (c * c).sum(axis=0, dtype=np.int64)
    power = power.reshape(nf, ns, 2, 2).sum(axis=3)  # [f, s, p]
    d = np.arange(ns) * (np.arange(ns) + 1) // 2 + np.arange(ns)
    autos = tri[:, d][:, :, [0, 3], 0]  # [f, s, (xx, yy)] real parts
    assert np.array_equal(autos, power)
    assert not tri[:, d][:, :, [0, 3], 1].any()  # autocorrelations are real


def _limit_dump_check(eng, tri_d, nf, ns, ref_chans, chans, compare):
    """Device triangular dump of eng into tri_d, then each of `chans` copied
    back and compared with ref_chans[c] by `compare`."""
    X.check(X.core().xgpuAmdDump(eng.handle, tri_d.ptr, X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR))
    nb = ns * (ns + 1) // 2
    for c, f in enumerate(chans):
        dt = np.int32 if eng.out_dtype == np.int32 else np.float32
        got = tri_d.download(dt, nb * 8, offset=f * nb * 8 * 4)
        compare(got, ref_chans[c], f)


@pytest.mark.parametrize("nf", [8191, 8192])
def test_i8_largest_blocks_at_the_offset_limit(nf):
    """NSTATION=512 x 2 pol, NTIME=64: NFREQUENCY=8191 is the largest block the
    real-Gram kernels take (128 * tstride just below 2^31 - 4096: their in-stage
    byte offsets are 32-bit), 8192 the smallest that falls back to the
    complex-plane kernel (64-bit offsets).  About 34 GB of int32 output.  At
    8191 every int8 kernel variant runs (the default DMA kernel, the 64-sample
    DMA and 16x16x64 variants, the register-staged kernels, the K-interleaved
    and the complex-plane kernel).  Exact on every visibility of the first,
    middle and last channel, and (default kernel) on sampled cells of every
    64th channel, through a device triangular dump of which only those parts
    are copied back."""
    ns, nt = 512, 64
    assert (128 * nf * ns * 4 < (1 << 31) - 4096) == (nf == 8191)
    flush(f"[i8] offset limit NS={ns} NF={nf} NT={nt}")
    nbytes = nt * nf * ns * 4
    buf = X.DeviceBuffer(nbytes)
    buf.fill_uniform_int8(nf, -128, 127)
    x = buf.download(np.int8)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    tri_d = X.DeviceBuffer(2 * eng.tri_len * 4)
    chans = [0, nf // 2, nf - 1]
    whole = oracle.xengine_i64_channels(x, ns, nf, nt, chans)

    def exact(got, ref, f):
        assert np.array_equal(got.reshape(-1, 4, 2).astype(np.int64), ref), f

    variants = [(0, 0, 0, 0)]
    if nf == 8191:
        variants += [(2, 2, 64, 0), (2, 2, 64, 2), (2, 1, 64, 0), (2, 1, 128, 0), (3, 0, 0, 0),
                     (1, 0, 0, 0)]
    for kernel, staging, stage, mfma in variants:
        flush(f"[i8] offset limit NF={nf} kernel={kernel} staging={staging} stage={stage} mfma={mfma}")
        eng.set_tuning(kernel=kernel, staging=staging, stage=stage, mfma=mfma)
        eng.clear()
        eng.correlate(buf, nt)
        _limit_dump_check(eng, tri_d, nf, ns, whole, chans, exact)
    eng.set_tuning()
    eng.clear()
    eng.correlate(buf, nt)
    X.check(X.core().xgpuAmdDump(eng.handle, tri_d.ptr, X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR))
    eng.close()
    buf.free()
    nb = ns * (ns + 1) // 2
    rng = np.random.default_rng(nf)
    fij = [(f, int(i), int(rng.integers(i + 1))) for f in range(0, nf, 64)
           for i in rng.integers(0, ns, 2)]
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, i, j) in enumerate(fij):
        k = f * nb + i * (i + 1) // 2 + j
        got = tri_d.download(np.int32, 8, offset=k * 32).reshape(4, 2)
        assert np.array_equal(got.astype(np.int64), ref[n]), (f, i, j)
    tri_d.free()


def test_i8_empty_blocks():
    """ntime = 0: a plain call leaves the integration as it was; in a batch,
    an empty job leaves its engine's integration as it was (zeros after a
    clear) and the other jobs are exact."""
    flush("[i8] empty blocks")
    xa = oracle.fill_uniform(128 * 2 * 64 * 4, 61, -128, 127)
    xb = oracle.fill_uniform(200 * 3 * 96 * 4, 62, -128, 127)
    ra, rb = oracle.xengine_i64(xa, 64, 2, 128), oracle.xengine_i64(xb, 96, 3, 200)
    ba, bb = X.DeviceBuffer.from_array(xa), X.DeviceBuffer.from_array(xb)
    a = X.Engine(64, 2, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    b = X.Engine(96, 3, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    a.correlate(ba, 128)
    a.correlate(ba, 0)
    assert np.array_equal(a.dump(wide=True), ra)
    X.correlate_batch([(a, ba, 0), (b, bb, 200)])
    assert np.array_equal(a.dump(wide=True), ra)
    assert np.array_equal(b.dump(wide=True), rb)
    a.clear()
    X.correlate_batch([(a, ba, 0), (b, bb, 0)])
    assert not a.dump(wide=True).any()
    assert np.array_equal(b.dump(wide=True), rb)
    a.correlate(ba, 128)
    assert np.array_equal(a.dump(wide=True), ra)
    a.close()
    b.close()
    ba.free()
    bb.free()


def test_i8_8192_stations():
    """NSTATION=8192 x 2 pol (16384 inputs, 8256 tiles per channel), NFREQUENCY=8,
    NTIME=256, register-tile order (8.6 GB of int32 output): every visibility
    of the last channel and every autocorrelation of the first against the
    int64 oracle and the input power, and sampled cells of every channel,
    through a device triangular dump."""
    ns, nf, nt = 8192, 8, 256
    flush(f"[i8] NS={ns} NF={nf} NT={nt}")
    buf = X.DeviceBuffer(nt * nf * ns * 4)
    buf.fill_uniform_int8(8192, -128, 127)
    x = buf.download(np.int8)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    eng.correlate(buf, nt)
    buf.free()
    tri_d = X.DeviceBuffer(2 * eng.tri_len * 4)
    X.check(X.core().xgpuAmdDump(eng.handle, tri_d.ptr, X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR))
    eng.close()
    nb = ns * (ns + 1) // 2
    last = tri_d.download(np.int32, nb * 8, offset=(nf - 1) * nb * 32).reshape(nb, 4, 2)
    flush("[i8] NS=8192: whole last channel against the oracle")
    ref = oracle.xengine_i64_channels(x, ns, nf, nt, [nf - 1])[0]
    assert np.array_equal(last.astype(np.int64), ref)
    del last, ref
    first = tri_d.download(np.int32, nb * 8, offset=0).reshape(nb, 4, 2)
    xt = x.reshape(nt, nf, ns, 2, 2)[:, 0].astype(np.int32)
    power = (xt * xt).sum(axis=(0, 3), dtype=np.int64)  # [s, p]
    d = np.arange(ns) * (np.arange(ns) + 1) // 2 + np.arange(ns)
    assert np.array_equal(first[d][:, [0, 3], 0].astype(np.int64), power)
    del first
    rng = np.random.default_rng(8192)
    fij = [(f, int(i), int(rng.integers(i + 1))) for f in range(nf) for i in rng.integers(0, ns, 32)]
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, i, j) in enumerate(fij):
        k = f * nb + i * (i + 1) // 2 + j
        got = tri_d.download(np.int32, 8, offset=k * 32).reshape(4, 2)
        assert np.array_equal(got.astype(np.int64), ref[n]), (f, i, j)
    tri_d.free()


def test_i8_largest_station_count():
    """NSTATION=32640 x 2 pol, the largest station count an engine takes (510
    row blocks, 130,305 tiles in one channel), NFREQUENCY=1, NTIME=64 (17 GB of
    int32 output): the whole last row of baselines (i = 32639, every j) and
    sampled cells against the int64 oracle, through a device triangular dump."""
    ns, nf, nt = 32640, 1, 64
    flush(f"[i8] NS={ns} NF={nf} NT={nt}")
    buf = X.DeviceBuffer(nt * nf * ns * 4)
    buf.fill_uniform_int8(ns, -128, 127)
    x = buf.download(np.int8)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    eng.correlate(buf, nt)
    buf.free()
    tri_d = X.DeviceBuffer(2 * eng.tri_len * 4)
    X.check(X.core().xgpuAmdDump(eng.handle, tri_d.ptr, X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR))
    eng.close()
    i = ns - 1
    row = tri_d.download(np.int32, (i + 1) * 8, offset=i * (i + 1) // 2 * 32).reshape(i + 1, 4, 2)
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, [(0, i, j) for j in range(i + 1)])
    assert np.array_equal(row.astype(np.int64), ref)
    rng = np.random.default_rng(ns)
    fij = [(0, int(a), int(rng.integers(a + 1))) for a in rng.integers(0, ns, 300)]
    fij += [(0, int(a), int(a)) for a in rng.integers(0, ns, 100)]
    ref = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, a, b) in enumerate(fij):
        k = a * (a + 1) // 2 + b
        got = tri_d.download(np.int32, 8, offset=k * 32).reshape(4, 2)
        assert np.array_equal(got.astype(np.int64), ref[n]), (a, b)
    tri_d.free()


def test_f16_largest_blocks_at_the_offset_limit():
    """FP16 at NSTATION=512 x 2 pol, NFREQUENCY=4095, NTIME=64: the largest
    block the FP16 real-Gram kernels take (tstride 16.8 MB; 17 GB of float
    output), for the DMA kernels rgdh<16,8> (default) and rgdh<32,4> and the
    register-staged rgh<32> and rgh<64>: the first, middle and last channel
    against the FP64 oracle within the fp32 accumulation bound."""
    ns, nf, nt = 512, 4095, 64
    assert 128 * nf * ns * 8 < (1 << 31) - 4096
    flush(f"[f16] offset limit NS={ns} NF={nf} NT={nt}")
    # 64 sample rows of complex Gaussian fp16 values (256 KiB) tiled over the
    # 1.07 GB block: sample row (t, f) gets row (t nf + f) mod 64
    base = f16_input(64, 1, ns, 91).reshape(-1)
    assert (nt * nf * ns * 4) % base.size == 0
    x16 = np.tile(base, nt * nf * ns * 4 // base.size).reshape(nt, nf, ns, 2, 2)
    buf = X.DeviceBuffer.from_array(x16)
    chans = [0, nf // 2, nf - 1]
    ref = oracle.xengine_f64_f16(np.ascontiguousarray(x16[:, chans]), ns, len(chans), nt)
    ref = ref.reshape(len(chans), -1)
    del x16
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    tri_d = X.DeviceBuffer(2 * eng.tri_len * 4)

    def close(got, r, f):
        assert_f16_close(got, r, ns, 1, nt)

    for kernel, staging, stage in [(0, 0, 0), (2, 2, 32), (2, 1, 32), (2, 1, 64)]:
        flush(f"[f16] offset limit kernel={kernel} staging={staging} stage={stage}")
        eng.set_tuning(kernel=kernel, staging=staging, stage=stage)
        eng.clear()
        eng.correlate(buf, nt)
        _limit_dump_check(eng, tri_d, nf, ns, ref, chans, close)
    eng.close()
    buf.free()
    tri_d.free()


def test_f16_cfg4_shape_errors_vs_f64():
    """NSTATION=256 x 2 pol, NFREQUENCY=128, NTIME=2048, fp16 complex
    N(0, 16^2) plus a delayed tone 4 e^{i(2 pi f tau_s / NF + w t)} (the cfg4
    generator).  Against the FP64 oracle of the same fp16 inputs on 4
    channels: max |dV| <= 1e-5 max|V|, relative RMS <= 1e-6, and normalized
    error |dV_ij| / sqrt(V_ii V_jj) <= (NT/16) 2^-24.  The fp16 products are
    exact in fp32 and each MFMA adds 16 of them to an fp32 accumulator, so the
    accumulation error is at most (NT/16) u sum_t |x_i x_j| with u = 2^-24, and
    sum_t |x_i x_j| <= sqrt(V_ii V_jj) (Cauchy-Schwarz).  The measured figures
    are printed (bench.py reports the same statistics)."""
    ns, nf, nt = 256, 128, 2048
    flush("[f16] cfg4 shape")
    x = W.cfg4_input(ns, nf, nt, seed=4)
    buf = X.DeviceBuffer.from_array(x)
    eng = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
    eng.correlate(buf, nt)
    got = eng.dump().reshape(nf, -1)
    eng.close()
    buf.free()
    chans = [0, 37, 90, 127]
    ref = oracle.xengine_f64_f16(np.ascontiguousarray(x[:, chans]), ns, len(chans), nt)
    st = W.vis_error_stats(got[chans], ref, ns)
    flush(f"[f16] cfg4 errors vs f64: {st}")
    assert st["max_abs_err"] <= 1e-5 * st["max_abs_vis"], st
    assert st["rms_rel_err"] <= 1e-6, st
    assert st["max_norm_err"] <= (nt / 16) * 2.0 ** -24, st


def test_f16_against_fp32_plain_on_int8_valued_cfg4():
    """The cfg4 generator rounded to int8 (values exact in fp16 and fp32), fed
    to the FP32 plain kernel as int8 and to the FP16 MFMA kernel as fp16:
    both against the exact int64 oracle on 4 channels, and against each
    other.  Sums reach 6.6e7 > 2^24, so both round.  Bounds, u = 2^-24,
    per polarization pair normalized by sqrt(V_ii V_jj):
      FP16 MFMA (16 products per fp32 rounding step):  (NT/16) u,
      FP32 plain (one rounding per FMA, 2 NT FMAs per component): 2 NT u,
    and their difference at most the sum of the two."""
    ns, nf, nt = 256, 128, 2048
    flush("[f16/f32] cfg4 int8-valued")
    x = W.cfg4_input(ns, nf, nt, seed=4, dtype=np.int8)
    chans = [0, 64, 127]
    exact = oracle.xengine_i64_channels(x, ns, nf, nt, chans).reshape(len(chans), -1)
    b8 = X.DeviceBuffer.from_array(x)
    bh = X.DeviceBuffer.from_array(x.astype(np.float16))
    e32 = X.Engine(ns, nf, compute=X.XGPU_FLOAT32, order=X.TRIANGULAR_ORDER)
    e32.correlate(b8, nt)
    g32 = e32.dump().reshape(nf, -1)[chans].astype(np.float64)
    e32.close()
    e16 = X.Engine(ns, nf, compute=X.XGPU_FLOAT16, order=X.TRIANGULAR_ORDER)
    e16.correlate(bh, nt)
    g16 = e16.dump().reshape(nf, -1)[chans].astype(np.float64)
    e16.close()
    b8.free()
    bh.free()
    s16 = W.vis_error_stats(g16, exact, ns)
    s32 = W.vis_error_stats(g32, exact, ns)
    d = W.vis_error_stats(g16, g32, ns)
    flush(f"[f16/f32] fp16 vs exact {s16}; fp32 plain vs exact {s32}; fp16 vs fp32 plain {d}")
    u = 2.0 ** -24
    assert s16["max_norm_err"] <= (nt / 16) * u, s16
    assert s32["max_norm_err"] <= 2 * nt * u, s32
    assert d["max_norm_err"] <= (nt / 16 + 2 * nt) * u, d
    assert s16["rms_rel_err"] <= 1e-6 and s32["rms_rel_err"] <= 1e-5, (s16, s32)


# ------------------------------------------------------------ shape sweep
# Seeded random shapes over every kernel variant (tests/shape_sweep.py); the
# loads and stores of each case are modelled on the CPU first
# (test_dma_model.py::test_shape_sweep_in_bounds).


@pytest.mark.parametrize("case", [pytest.param(c, id=S.case_id(c)) for c in S.I8_SWEEP])
def test_i8_shape_sweep(case):
    """Bit-exact against the int64 oracle, 32-bit and 64-bit dumps, whichever
    kernel the engine launches for the case's tuning and each call's input
    (ragged tiles, time tails, accumulating calls, unaligned later calls)."""
    ns, nf, nt, order, calls, t = case
    x = oracle.fill_uniform(nt * nf * ns * 4, 31 * ns + nt, -128, 127)
    raw, wide = run_i8(ns, nf, nt, x, ORDERS[order], calls, **t)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    assert np.array_equal(to_tri(None, raw, ns, nf, ORDERS[order]).astype(np.int64), ref)
    assert np.array_equal(to_tri(None, wide, ns, nf, ORDERS[order]), ref)


@pytest.mark.parametrize("case", [pytest.param(c, id=S.case_id(c)) for c in S.F16_SWEEP])
def test_f16_shape_sweep(case):
    """Against the FP64 oracle of the same fp16 input, with the bounds of
    assert_f16_close; the 64-bit dump equals the 32-bit one."""
    ns, nf, nt, order, calls, t = case
    x16 = f16_input(nt, nf, ns, 17 * ns + nt)
    got, wide = run_f16(ns, nf, nt, x16, ORDERS[order], calls, **t)
    ref = oracle.xengine_f64_f16(x16, ns, nf, nt)
    assert_f16_close(got, ref, ns, nf, nt, calls=calls)
    assert np.array_equal(wide, got.astype(np.float64))
