"""GPU tests of the drop-in surface around the engine: the cuda_correlator
driver (the reference's only end-to-end check, reference
src/cuda_correlator.c:184-190 and :247-253), input_offset / output_offset
ring-buffer use through xgpu.h (reference src/xgpu.h:103-109,
src/cuda_xengine.cu:590, :658), and the device converters (swizzle, extract,
reference src/cpu_util.c:235-279)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import _lib
from conftest import flush

pytestmark = pytest.mark.gpu

LIB = _lib.LIB_DIR


def run_tool(name, *args, timeout=120):
    flush(f"[tools] {name} {' '.join(args)}")
    r = subprocess.run([os.path.join(LIB, name), *args], capture_output=True, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("tool", ["cuda_correlator", "cuda_correlator_dp4a"])
@pytest.mark.parametrize("caller_buffers", [False, True])
def test_cuda_correlator_checks_against_cpu_engine(tool, caller_buffers):
    """One dumping call, checked by xgpuCheckResult against xgpuOmpXengine on
    the xgpuRandomComplex data (the DP4A build swizzles its input first)."""
    rc, out = run_tool(tool, *(["-r"] if caller_buffers else []))
    assert rc == 0, out
    assert "Correlating 256 stations with 10 channels and integration length 1024" in out
    assert "Calling CPU X-Engine" in out and "Calling GPU X-Engine" in out
    assert "Outer product summation successful" in out, out
    assert "Elapsed time" in out and "Theoretical BW_max" in out


@pytest.mark.parametrize("args", [["-c", "3", "-o", "2"], ["-c", "2", "-C", "3", "-o", "0"],
                                  ["-c", "2", "-o", "3", "-f", "3"]])
def test_cuda_correlator_timing_modes(args):
    """Several calls per integration and several integrations: no check (as in
    the reference), one pair of rate lines per integration plus a total."""
    rc, out = run_tool("cuda_correlator", *args)
    assert rc == 0, out
    nint = int(args[args.index("-C") + 1]) if "-C" in args else 1
    lines = out.count("Elapsed time")
    assert lines == nint + (1 if nint > 1 else 0), out
    assert "Calling CPU X-Engine" not in out


def test_cuda_correlator_usage():
    rc, out = run_tool("cuda_correlator", "-h")
    assert rc == 0 and "-c calls" in out
    rc, out = run_tool("cuda_correlator", "-c", "0")
    assert rc != 0 and "count must be positive" in out


def test_xgpu_h_input_and_output_offsets():
    """Caller-allocated, oversized host buffers used as rings: each call reads
    NTIME samples at array_h + input_offset and dumps matLength values at
    matrix_h + output_offset; nothing outside those windows is written."""
    lib = X.variant("libxgpu_ns16_tri")
    info = _lib.XGPUInfo()
    lib.xgpuInfo(C.byref(info))
    ns, nf, nt = info.nstation, info.nfrequency, info.ntime
    vec, mat = info.vecLength, info.matLength
    nblk = 3
    ring = np.zeros(2 * (nblk * vec + 5), dtype=np.int8)       # ComplexInput = 2 bytes
    out = np.full(2 * (2 * mat + 7), -1.0, dtype=np.float32)    # Complex = 2 floats
    blocks = [oracle.fill_uniform(2 * vec, 300 + b, -127, 127) for b in range(nblk)]
    ctx = _lib.XGPUContext()
    ctx.array_h, ctx.matrix_h = ring.ctypes.data, out.ctypes.data
    ctx.array_len, ctx.matrix_len = nblk * vec + 5, 2 * mat + 7
    flush("[tools] xgpu.h offsets")
    assert lib.xgpuInit(C.byref(ctx), 0) == 0
    assert ctx.input_offset == 0 and ctx.output_offset == 0
    in_off = [5, 5 + vec, 5 + 2 * vec]  # ComplexInput units, odd start
    for b in range(nblk):
        ring[2 * in_off[b]: 2 * (in_off[b] + vec)] = blocks[b]
    # integration 1: blocks 0 and 1, dumped at output_offset 7
    for b, op in ((0, X.SYNCOP_SYNC_TRANSFER), (1, X.SYNCOP_DUMP)):
        ctx.input_offset = in_off[b]
        ctx.output_offset = 7
        assert lib.xgpuCudaXengine(C.byref(ctx), op) == 0
    # integration 2: block 2 alone, dumped right after it
    assert lib.xgpuClearDeviceIntegrationBuffer(C.byref(ctx)) == 0
    ctx.input_offset = in_off[2]
    ctx.output_offset = 7 + mat
    assert lib.xgpuCudaXengine(C.byref(ctx), X.SYNCOP_DUMP) == 0
    ref1 = oracle.xengine_f32(blocks[0], ns, nf, nt) + oracle.xengine_f32(blocks[1], ns, nf, nt)
    ref2 = oracle.xengine_f32(blocks[2], ns, nf, nt)
    assert np.array_equal(out[14:14 + 2 * mat], ref1)
    assert np.array_equal(out[14 + 2 * mat:14 + 4 * mat], ref2)
    assert np.all(out[:14] == -1.0)
    lib.xgpuFree(C.byref(ctx))
    assert ctx.array_h is None and ctx.matrix_h is None


# ------------------------------------------------------------------ converters
@pytest.mark.parametrize("ns,nf,nt", [(16, 2, 64), (6, 3, 20), (256, 10, 1024)])
def test_device_swizzle_matches_oracle_and_inverts(ns, nf, nt):
    x = oracle.fill_uniform(nt * nf * ns * 4, 400 + ns, -128, 127)
    src = X.DeviceBuffer.from_array(x)
    sw = X.DeviceBuffer(x.nbytes)
    back = X.DeviceBuffer(x.nbytes)
    flush(f"[tools] swizzle ns={ns} nf={nf} nt={nt}")
    X.swizzle_input(sw, src, ns, nf, nt)
    assert np.array_equal(sw.download(np.int8), oracle.swizzle(x, ns, nf, nt))
    X.swizzle_input(back, sw, ns, nf, nt, unswizzle=True)
    assert np.array_equal(back.download(np.int8), x)
    for b in (src, sw, back):
        b.free()


def test_swizzled_input_engine_equals_plain():
    """An engine told its input is corner-turned (time_swizzle 4) gives the
    plain input's exact result, on the device path, in a batch and on the
    host-buffer path (chunked transfers, each chunk unswizzled on the device)."""
    ns, nf, nt = 64, 3, 512
    x = oracle.fill_uniform(nt * nf * ns * 4, 410, -128, 127)
    ref = oracle.xengine_i64(x, ns, nf, nt)
    xs = oracle.swizzle(x, ns, nf, nt)
    flush("[tools] swizzled-input engine")
    buf = X.DeviceBuffer.from_array(xs)
    eng = X.Engine(ns, nf, nt, 128, compute=X.XGPU_INT8, order=X.TRIANGULAR_ORDER,
                   time_swizzle=4)
    eng.correlate(buf, nt)
    assert np.array_equal(eng.dump(wide=True), ref)
    eng.clear()
    X.correlate_batch([(eng, buf, nt)])
    assert np.array_equal(eng.dump(wide=True), ref)
    eng.clear()
    got = eng.correlate_host(xs)
    assert np.array_equal(got.astype(np.int64), ref)
    eng.close()
    buf.free()


@pytest.mark.parametrize("ns,nf", [(24, 3), (1, 1), (16, 2), (37, 2), (130, 1)])
@pytest.mark.parametrize("kind,dtype", [(X.ELEM_INT32, np.int32), (X.ELEM_INT64, np.int64),
                                        (X.ELEM_FLOAT32, np.float32),
                                        (X.ELEM_FLOAT64, np.float64)])
def test_device_extract_matches_oracle(kind, dtype, ns, nf):
    """Device xgpuExtractMatrix against the oracle's (pinned to the
    reference's golden vectors), on values every element type holds exactly:
    tiles of 16 x 16 stations, whole (16), partial (24, 37, 130) and a single
    station."""
    n = 2 * oracle.tri_len(ns, nf)
    vals = oracle.fill_uniform(n, 420 + kind, -128, 127).astype(np.int64) * 4099
    packed = vals.astype(dtype)
    ref = oracle.extract(vals.astype(np.float32), ns, nf)
    src = X.DeviceBuffer.from_array(packed)
    dst = X.DeviceBuffer(ref.size * np.dtype(dtype).itemsize)
    flush(f"[tools] extract kind={kind}")
    X.extract_matrix(dst, src, ns, nf, kind)
    got = dst.download(dtype).astype(np.float32)
    assert np.array_equal(got, ref)
    src.free()
    dst.free()


def test_device_extract_of_a_dump_is_hermitian():
    """Extract of a device triangular dump: M[j][i] = conj(M[i][j]) for every
    pair, the lower half equal to the dump."""
    ns, nf, nt = 40, 2, 128
    x = oracle.fill_uniform(nt * nf * ns * 4, 430, -128, 127)
    eng = X.Engine(ns, nf, compute=X.XGPU_INT8, order=X.REGISTER_TILE_TRIANGULAR_ORDER)
    buf = X.DeviceBuffer.from_array(x)
    eng.correlate(buf, nt)
    tri = X.DeviceBuffer(2 * eng.tri_len * 8)
    X.check(X.core().xgpuAmdDump(eng.handle, tri.ptr,
                                 X.XGPU_DUMP_DEVICE | X.XGPU_DUMP_TRIANGULAR | X.XGPU_DUMP_WIDE))
    full = X.DeviceBuffer(nf * ns * ns * 4 * 2 * 8)
    flush("[tools] extract of a dump")
    X.extract_matrix(full, tri, ns, nf, X.ELEM_INT64)
    m = full.download(np.int64).reshape(nf, ns, ns, 2, 2, 2)
    ref = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 2, 2, 2)
    for i in range(ns):
        for j in range(i + 1):
            k = i * (i + 1) // 2 + j
            assert np.array_equal(m[:, i, j], ref[:, k])
            assert np.array_equal(m[:, j, i, :, :, 0], ref[:, k, :, :, 0].transpose(0, 2, 1))
            assert np.array_equal(m[:, j, i, :, :, 1], -ref[:, k, :, :, 1].transpose(0, 2, 1))
    for b in (buf, tri, full):
        b.free()
    eng.close()
