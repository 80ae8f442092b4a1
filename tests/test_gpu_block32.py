"""GPU parity of the COMPLEX_BLOCK_SIZE=32 input layout.

In that layout (reference src/omp_xengine.c:64-78) each block of 32
ComplexInput entries holds the real parts of 16 stations' (X, Y) followed by
their imaginary parts.  The engine converts it on the device (a 64-byte block
permutation) before the kernels.  Checks: the device path for every int8
kernel variant and the FP32 plain kernel against the oracle of the unblocked
input (bit-exact), the xgpu.h host path of the block-32 builds against the
reference's own block-32 CPU engine output (golden vectors), and a ragged batch
mixing a block-32 engine with a plain one.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import _lib
from conftest import flush, golden

pytestmark = pytest.mark.gpu

ORDERS = {
    "regtile": X.REGISTER_TILE_TRIANGULAR_ORDER,
    "tri": X.TRIANGULAR_ORDER,
    "realimag": X.REAL_IMAG_TRIANGULAR_ORDER,
}


def to_tri(raw, ns, nf, order):
    return raw if order == X.TRIANGULAR_ORDER else oracle.reorder(order, raw, ns, nf)


def run_block32(ns, nf, nt, xb, compute, order, calls=1, **tuning):
    flush(f"[block32] ns={ns} nf={nf} nt={nt} compute={compute} order={order} "
          f"calls={calls} {tuning}")
    eng = X.Engine(ns, nf, compute=compute, order=order, complex_block_size=32)
    if tuning:
        eng.set_tuning(**tuning)
    buf = X.DeviceBuffer.from_array(xb)
    per = nt // calls
    step = per * nf * ns * 4
    for c in range(calls):
        n = per if c < calls - 1 else nt - per * (calls - 1)
        eng.correlate(buf.ptr + c * step, n)
    wide = eng.dump(wide=True)
    eng.close()
    buf.free()
    return to_tri(wide, ns, nf, order)


@pytest.mark.parametrize("ns,nf,nt,order,calls", [
    (32, 3, 256, "tri", 1),
    (48, 2, 300, "regtile", 2),     # 96 inputs: one partial tile per side
    (144, 2, 100, "realimag", 2),   # 288 inputs: three tiles per side
])
@pytest.mark.parametrize("kernel,staging,stage", [(1, 0, 0), (2, 1, 64), (2, 2, 32), (3, 0, 0)])
def test_block32_int8_exact(ns, nf, nt, order, calls, kernel, staging, stage):
    xb = oracle.fill_uniform(nt * nf * ns * 4, 60 + ns, -128, 127)
    got = run_block32(ns, nf, nt, xb, X.XGPU_INT8, ORDERS[order], calls,
                      kernel=kernel, staging=staging, stage=stage)
    ref = oracle.xengine_i64(oracle.unblock32(xb), ns, nf, nt)
    assert np.array_equal(got, ref)
    # the layout matters: the same bytes read as plain input differ
    assert not np.array_equal(got, oracle.xengine_i64(xb, ns, nf, nt))


def test_block32_fp32_plain_matches_float_oracle():
    """FP32 plain kernel on blocked input: bit-identical to the float CPU
    engine of the unblocked input (every partial sum below 2^24 is exact)."""
    ns, nf, nt = 32, 3, 256
    xb = oracle.fill_uniform(nt * nf * ns * 4, 71, -127, 127)
    got = run_block32(ns, nf, nt, xb, X.XGPU_FLOAT32, X.TRIANGULAR_ORDER)
    ref = oracle.xengine_f32(oracle.unblock32(xb), ns, nf, nt)
    assert np.array_equal(got, ref.astype(np.float64))


@pytest.mark.parametrize("variant,tag,ctype", [
    ("libxgpu_ns16_cb32", "ns16_nf2_nt64_cb32", C.c_float),
    ("libxgpu_ns32_cb32_dp4a", "ns32_nf3_nt256_cb32_dp4a", C.c_int32)])
def test_block32_through_xgpu_h_matches_reference(meta, variant, tag, ctype):
    """xgpuInit -> xgpuCudaXengine(SYNCOP_DUMP) -> xgpuReorderMatrix of a
    COMPLEX_BLOCK_SIZE=32 build equals the reference's CPU X-engine built the
    same way (golden vectors; sums below 2^24, so the float build is exact)."""
    m = meta[tag]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    flush(f"[block32] {variant} via xgpu.h")
    lib = X.variant(variant)
    ctx = _lib.XGPUContext()
    assert lib.xgpuInit(C.byref(ctx), 0) == 0
    info = _lib.XGPUInfo()
    lib.xgpuInfo(C.byref(info))
    assert (info.nstation, info.nfrequency, info.ntime, info.complex_block_size) == (ns, nf, nt, 32)
    xb = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    if "dp4a" in variant:
        # DP4A builds read the corner-turned input, as the reference's do
        # (src/cuda_correlator.c:172-173)
        lib.xgpuSwizzleInput(ctx.array_h, xb.ctypes.data)
    else:
        C.memmove(ctx.array_h, xb.ctypes.data, xb.nbytes)
    assert lib.xgpuCudaXengine(C.byref(ctx), X.SYNCOP_DUMP) == 0
    mat = np.ctypeslib.as_array((ctype * (2 * info.matLength)).from_address(ctx.matrix_h))
    lib.xgpuReorderMatrix(ctx.matrix_h)
    got = mat[: 2 * info.triLength].copy()
    ref = golden(f"ref_{tag}.npz")["omp"]
    assert np.array_equal(got, ref)
    lib.xgpuFree(C.byref(ctx))


def test_block32_in_a_ragged_batch():
    """One launch for a block-32 engine and a plain engine: each equals its
    oracle; the batch accumulates on a second call."""
    shapes = [(32, 3, 256, 32), (40, 2, 200, 1)]
    engines, bufs, refs, ins = [], [], [], []
    for k, (ns, nf, nt, cbs) in enumerate(shapes):
        x = oracle.fill_uniform(nt * nf * ns * 4, 80 + k, -128, 127)
        engines.append(X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER,
                                complex_block_size=cbs))
        bufs.append(X.DeviceBuffer.from_array(x))
        plain = oracle.unblock32(x) if cbs == 32 else x
        refs.append(oracle.xengine_i64(plain, ns, nf, nt))
    flush("[block32] ragged batch")
    items = [(e, b, s[2]) for e, b, s in zip(engines, bufs, shapes)]
    X.correlate_batch(items)
    for e, r in zip(engines, refs):
        assert np.array_equal(e.dump().astype(np.int64), r)
    X.correlate_batch(items)
    for e, r in zip(engines, refs):
        assert np.array_equal(e.dump(wide=True), 2 * r)
        e.close()
    for b in bufs:
        b.free()
