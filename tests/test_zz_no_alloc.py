"""GPU test of xgpuAmdCorrelateNoAlloc (the entry point that allocates
nothing: its input conversion runs through the device's static scratch)
against xgpuAmdCorrelate, called through the C-ABI.  Bit-exact on two calls
in a row, the second with a smaller input; the first input is larger than
the 64 MiB scratch, so it is converted and correlated in chunks."""
import ctypes as C

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from conftest import flush

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _dump(eng):
    return eng.dump(wide=True, triangular=True)


@pytest.mark.parametrize("layout", ["swizzle", "block32"])
def test_no_alloc_equals_correlate(layout):
    ns, nf = 256, 64                      # 64 KiB per sample: 1024 samples fill the scratch
    tstride = nf * ns * 4
    sizes = (2500, 300)                   # > scratch (3 chunks), then smaller
    kw = {"time_swizzle": 4} if layout == "swizzle" else {"complex_block_size": 32}
    a = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER, **kw)
    b = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER, **kw)
    assert sizes[0] * tstride > SCRATCH
    for k, nt in enumerate(sizes):
        flush(f"[no-alloc] {layout} call {k} ntime={nt}")
        if layout == "swizzle":
            plain = oracle.fill_uniform(nt * tstride, 500 + k, -128, 127)
            x = oracle.swizzle(plain, ns, nf, nt)
        else:
            x = oracle.fill_uniform(nt * tstride, 500 + k, -128, 127)
            plain = oracle.unblock32(x)
        buf = X.DeviceBuffer.from_array(x)
        a.clear()
        b.clear()
        X.check(X.core().xgpuAmdCorrelateNoAlloc(a.handle, buf.ptr, nt))
        X.check(X.core().xgpuAmdCorrelate(b.handle, buf.ptr, nt))
        da, db = _dump(a), _dump(b)
        assert np.array_equal(da, db), (layout, k)
        # and against the oracle on one channel of the plain input
        ref = oracle.xengine_i64_channels(plain, ns, nf, nt, [nf - 1])
        assert np.array_equal(da.reshape(nf, -1)[nf - 1], ref.reshape(-1)), (layout, k)
        buf.free()
    # accumulating: one more call into each integration keeps them equal
    buf = X.DeviceBuffer(sizes[1] * tstride)
    buf.fill_uniform_int8(600, -128, 127)
    X.check(X.core().xgpuAmdCorrelateNoAlloc(a.handle, buf.ptr, sizes[1]))
    X.check(X.core().xgpuAmdCorrelate(b.handle, buf.ptr, sizes[1]))
    assert np.array_equal(_dump(a), _dump(b))
    buf.free()
    a.close()
    b.close()


def test_no_alloc_plain_input_and_refusals():
    """Plain input needs no scratch; a call that would pass the int32 bound
    before the spill buffer exists is refused before any device work."""
    ns, nf, nt = 64, 4, 512
    x = oracle.fill_uniform(nt * nf * ns * 4, 610, -128, 127)
    buf = X.DeviceBuffer.from_array(x)
    e = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER)
    flush("[no-alloc] plain input")
    for _ in range(2):
        X.check(X.core().xgpuAmdCorrelateNoAlloc(e.handle, buf.ptr, nt))
    assert np.array_equal(e.dump(wide=True), 2 * oracle.xengine_i64(x, ns, nf, nt))
    big = X.I32_SAMPLE_BOUND
    e.clear()
    rc = X.core().xgpuAmdCorrelateNoAlloc(e.handle, buf.ptr, big + 1)
    assert rc == X._lib.XGPU_OUT_OF_MEMORY
    assert b"spill buffer" in X.core().xgpuAmdLastError()
    e.close()
    buf.free()
