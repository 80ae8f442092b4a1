"""xgpu_amd -- MI355X-native X-engine with the xGPU ``xgpu.h`` C API.

The compute runs in hand-written gfx950 HIP kernels behind a C-ABI
(``include/xgpu.h``, ``include/xgpu_amd.h``).  This package is the Python host
side: ctypes bindings plus small helpers for device buffers and engines.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (  # noqa: F401  (re-exported constants)
    REAL_IMAG_TRIANGULAR_ORDER, REGISTER_TILE_TRIANGULAR_ORDER, SYNCOP_DUMP, SYNCOP_NONE,
    SYNCOP_SYNC_COMPUTE, SYNCOP_SYNC_TRANSFER, TRIANGULAR_ORDER, XGPU_DUMP_DEVICE,
    XGPU_DUMP_TRIANGULAR, XGPU_DUMP_WIDE, XGPU_FLOAT16, XGPU_FLOAT32, XGPU_INT8, XGPUAmdConfig,
    XGPUContext, XGPUInfo, XgpuError, XgpuLibraryMissing, check, core, variant,
)

__all__ = ["Engine", "DeviceBuffer", "HostBuffer", "Server", "info", "cmac_per_integration",
           "correlate_batch"]

I32_SAMPLE_BOUND = 65535  # samples an int32 integration holds for any int8 input


def cmac_per_integration(nstation: int, nfreq: int, ntime: int, npol: int = 2) -> int:
    """Complex multiply-accumulates of one integration, NF*NT*NS(NS+1)/2*NPOL^2."""
    return nfreq * ntime * nstation * (nstation + 1) // 2 * npol * npol


def _config(nstation, nfrequency, ntime=0, ntimepipe=0, compute=XGPU_INT8,
            order=REGISTER_TILE_TRIANGULAR_ORDER, device=0, complex_block_size=1,
            time_swizzle=0) -> XGPUAmdConfig:
    c = XGPUAmdConfig()
    c.npol, c.nstation, c.nfrequency = 2, nstation, nfrequency
    c.ntime, c.ntimepipe = ntime, (ntimepipe or ntime)
    c.compute_type, c.matrix_order, c.device = compute, order, device
    c.complex_block_size = complex_block_size
    c.time_swizzle = time_swizzle
    return c


def info(nstation, nfrequency, ntime=0, ntimepipe=0, compute=XGPU_INT8,
         order=REGISTER_TILE_TRIANGULAR_ORDER, complex_block_size=1) -> XGPUInfo:
    out = XGPUInfo()
    cfg = _config(nstation, nfrequency, ntime, ntimepipe, compute, order,
                  complex_block_size=complex_block_size)
    check(core().xgpuAmdInfo(C.byref(cfg), C.byref(out)))
    return out


class DeviceBuffer:
    """A block of device memory owned by this object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(core().xgpuAmdMalloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(core().xgpuAmdMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, _lib.XGPU_H2D))

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else count
        assert 0 <= offset and offset + n * dt.itemsize <= self.nbytes
        out = np.empty(n, dtype=dt)
        check(core().xgpuAmdMemcpy(out.ctypes.data, self.ptr + offset, out.nbytes, _lib.XGPU_D2H))
        return out

    def fill_uniform_int8(self, seed: int, lo: int, hi: int, nbytes: int | None = None) -> None:
        n = self.nbytes if nbytes is None else nbytes
        check(core().xgpuAmdFillUniformInt8(self.ptr, n, seed, lo, hi))

    def free(self) -> None:
        if self.ptr:
            core().xgpuAmdFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One subarray's X-engine (runtime-sized), see include/xgpu_amd.h."""

    def __init__(self, nstation, nfrequency, ntime=0, ntimepipe=0, compute=XGPU_INT8,
                 order=REGISTER_TILE_TRIANGULAR_ORDER, device=0, complex_block_size=1,
                 time_swizzle=0):
        """complex_block_size: input layout, 1 = plain [t][f][s][p], 32 = the
        COMPLEX_BLOCK_SIZE=32 blocked layout (int8 input, nstation % 16 == 0);
        time_swizzle: 4 = input in the depth-4 time corner turn of
        xgpuSwizzleInput (the reference's DP4A input layout)."""
        self.cfg = _config(nstation, nfrequency, ntime, ntimepipe, compute, order, device,
                           complex_block_size, time_swizzle)
        self.info = XGPUInfo()
        check(core().xgpuAmdInfo(C.byref(self.cfg), C.byref(self.info)))
        h = C.c_void_p()
        check(core().xgpuAmdCreate(C.byref(h), C.byref(self.cfg)))
        self.handle = h.value
        self.compute = compute
        self.nstation, self.nfrequency = nstation, nfrequency

    # sizes
    @property
    def mat_len(self) -> int:
        return int(self.info.matLength)

    @property
    def tri_len(self) -> int:
        return int(self.info.triLength)

    @property
    def out_dtype(self):
        return np.int32 if self.compute == XGPU_INT8 else np.float32

    @property
    def wide_dtype(self):
        return np.int64 if self.compute == XGPU_INT8 else np.float64

    def clear(self) -> None:
        check(core().xgpuAmdClear(self.handle))

    def correlate(self, input_d, ntime: int) -> None:
        ptr = input_d.ptr if isinstance(input_d, DeviceBuffer) else int(input_d)
        check(core().xgpuAmdCorrelate(self.handle, ptr, ntime))

    def correlate_host(self, input_h: np.ndarray, sync_op: int = SYNCOP_DUMP):
        out = np.zeros(2 * self.mat_len, dtype=self.out_dtype)
        a = np.ascontiguousarray(input_h)
        check(core().xgpuAmdCorrelateHost(self.handle, a.ctypes.data, sync_op, out.ctypes.data))
        return out

    def dump(self, wide: bool = False, triangular: bool = False) -> np.ndarray:
        """Integration as a flat array: split planes (register tile / real-imag)
        or interleaved complex (triangular order or ``triangular=True``)."""
        n = 2 * (self.tri_len if triangular else self.mat_len)
        out = np.empty(n, dtype=self.wide_dtype if wide else self.out_dtype)
        flags = (_lib.XGPU_DUMP_WIDE if wide else 0) | (_lib.XGPU_DUMP_TRIANGULAR if triangular else 0)
        check(core().xgpuAmdDump(self.handle, out.ctypes.data, flags))
        return out

    def synchronize(self) -> None:
        check(core().xgpuAmdSynchronize(self.handle))

    @property
    def samples(self) -> int:
        return int(core().xgpuAmdSamples(self.handle))

    def time_correlate(self, input_d, ntime: int, warmup: int = 1, reps: int = 5) -> float:
        ms = C.c_float()
        ptr = input_d.ptr if isinstance(input_d, DeviceBuffer) else int(input_d)
        check(core().xgpuAmdTimeCorrelate(self.handle, ptr, ntime, warmup, reps, C.byref(ms)))
        return float(ms.value)

    _TUNING = ("wg_per_cu", "prefetch", "graph", "kernel", "stage", "staging", "ntimepipe", "mfma")

    def set_tuning(self, wg_per_cu: int = 0, prefetch: int = 0, graph: int = 0,
                   kernel: int = 0, stage: int = 0, staging: int = 0, ntimepipe: int = 0,
                   mfma: int = 0) -> None:
        """wg_per_cu: 1 or 2; prefetch / graph: 1 on, 2 off; kernel (int8 and
        fp16): 1 complex-plane, 2 real-Gram where it applies, 3 (int8)
        K-interleaved pair-tile kernel where the real-Gram one applies; staging
        (real-Gram): 1 registers, 2 global->LDS DMA; stage: samples per LDS
        stage (int8: 64 / 128 with register staging, 32 / 64 with DMA; fp16:
        32 / 64 with register staging, 16 / 32 with DMA); ntimepipe: host-path
        samples per transfer and launch; mfma (int8 DMA kernel): 1 32x32x32,
        2 16x16x64 tiles (0 = default everywhere)."""
        t = _lib.XGPUAmdTuning()
        t.wg_per_cu, t.prefetch, t.graph = wg_per_cu, prefetch, graph
        t.kernel, t.stage, t.staging = kernel, stage, staging
        t.ntimepipe, t.mfma = ntimepipe, mfma
        check(core().xgpuAmdSetTuning(self.handle, C.byref(t)))

    def get_tuning(self) -> dict:
        t = _lib.XGPUAmdTuning()
        check(core().xgpuAmdGetTuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k in self._TUNING}

    def autotune(self, input_d, ntime: int, warmup: int = 1, reps: int = 3,
                 cache_path: str | None = None):
        """Time every kernel tuning on this input (on a scratch engine: this
        engine's integration is left as it is); keep and return the fastest."""
        t = _lib.XGPUAmdTuning()
        ms = C.c_float()
        ptr = input_d.ptr if isinstance(input_d, DeviceBuffer) else int(input_d)
        check(core().xgpuAmdAutotune(self.handle, ptr, ntime, warmup, reps,
                                     cache_path.encode() if cache_path else None,
                                     C.byref(t), C.byref(ms)))
        return ({k: getattr(t, k) for k in self._TUNING}, float(ms.value))

    def autotune_host(self, input_h, warmup: int = 1, reps: int = 3,
                      cache_path: str | None = None):
        """Time the host-buffer path's time-pipe lengths x graph replay on this
        host input; keep and return the fastest (ms per call)."""
        t = _lib.XGPUAmdTuning()
        ms = C.c_float()
        ptr = input_h.ptr if isinstance(input_h, HostBuffer) else np.ascontiguousarray(input_h).ctypes.data
        check(core().xgpuAmdAutotuneHost(self.handle, ptr, warmup, reps,
                                         cache_path.encode() if cache_path else None,
                                         C.byref(t), C.byref(ms)))
        return ({k: getattr(t, k) for k in self._TUNING}, float(ms.value))

    def load_tuning(self, cache_path: str, ntime: int) -> bool:
        return core().xgpuAmdLoadTuning(self.handle, cache_path.encode(), ntime) == 0

    def close(self) -> None:
        if getattr(self, "handle", None):
            core().xgpuAmdDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_peak(device: int = 0, reps: int = 3) -> float:
    """Measured MFMA-only int8 throughput of the device, complex MACs per second."""
    v = C.c_double()
    check(core().xgpuAmdMfmaPeak(device, reps, C.byref(v)))
    return float(v.value)


PEAK_OPERAND_BYTES = 65536


def mfma_peak_data(operands, device: int = 0, reps: int = 3, compute: int = XGPU_INT8) -> float:
    """MFMA-only throughput (complex MACs per second) with operands taken from
    `operands` (a DeviceBuffer or device pointer holding PEAK_OPERAND_BYTES of
    int8 values, or fp16 values with compute=XGPU_FLOAT16): the MFMA ceiling
    for input of that distribution."""
    if isinstance(operands, DeviceBuffer):
        assert operands.nbytes >= PEAK_OPERAND_BYTES
        ptr = operands.ptr
    else:
        ptr = int(operands)
    v = C.c_double()
    check(core().xgpuAmdMfmaPeakData(device, ptr, compute, reps, C.byref(v)))
    return float(v.value)


def correlate_batch(items, stream=None) -> None:
    """Ragged batch in one persistent launch: items = [(engine, input_d, ntime)]."""
    arr = (_lib.XGPUAmdBatchItem * len(items))()
    for k, (eng, buf, nt) in enumerate(items):
        arr[k].engine = eng.handle
        arr[k].input_d = buf.ptr if isinstance(buf, DeviceBuffer) else int(buf)
        arr[k].ntime = nt
    check(core().xgpuAmdCorrelateBatch(arr, len(items), stream))


ELEM_INT32, ELEM_INT64, ELEM_FLOAT32, ELEM_FLOAT64 = 0, 1, 2, 3


def _ptr(b):
    return b.ptr if isinstance(b, (DeviceBuffer, HostBuffer)) else int(b)


def swizzle_input(dst, src, nstation: int, nfrequency: int, ntime: int, unswizzle=False,
                  stream=None) -> None:
    """Device xgpuSwizzleInput (or its inverse): plain [t][f][s][p][re,im]
    int8 input -> the depth-4 time corner turn [t/4][f][s][p][c][t%4]."""
    check(core().xgpuAmdSwizzleInput(_ptr(dst), _ptr(src), nstation, nfrequency, ntime,
                                     int(unswizzle), stream))
    check(core().xgpuAmdDeviceSynchronize())


def extract_matrix(dst, packed, nstation: int, nfrequency: int, elem: int, stream=None) -> None:
    """Device xgpuExtractMatrix: packed triangular -> full Hermitian matrix."""
    check(core().xgpuAmdExtractMatrix(_ptr(dst), _ptr(packed), nstation, nfrequency, elem, stream))
    check(core().xgpuAmdDeviceSynchronize())


FFT_NATURAL, FFT_PROTOTYPE = 0, 1
FFT_SIZES = (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192)


def fft(work, n: int, batch: int, inverse: bool = False, layout: int = FFT_NATURAL,
        stream=None) -> None:
    """In-place batched complex64 FFT of `batch` transforms of n points on the
    device (include/xgpu_fft.h); inverse is unnormalized."""
    ptr = work.ptr if isinstance(work, DeviceBuffer) else int(work)
    rc = _lib.fft_lib().xgpuAmdFft(ptr, n, batch, int(inverse), layout, stream)
    if rc:
        raise XgpuError(rc, f"xgpuAmdFft(n={n}, batch={batch}, layout={layout}) returned {rc}")


class HostBuffer:
    """Pinned host memory viewed as a numpy array (fast asynchronous copies)."""

    def __init__(self, nbytes: int, dtype=np.int8):
        p = C.c_void_p()
        check(core().xgpuAmdHostAlloc(C.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        dt = np.dtype(dtype)
        self.array = np.ctypeslib.as_array((C.c_byte * self.nbytes).from_address(self.ptr)).view(dt)

    def free(self):
        if self.ptr:
            self.array = None
            core().xgpuAmdHostFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching correlation server (include/xgpu_serve.h)."""

    def __init__(self, device=0, staging_bytes=0, max_batch=0, max_inflight=0):
        cfg = _lib.XGPUServerConfig(device, staging_bytes, max_batch, max_inflight)
        h = C.c_void_p()
        check(core().xgpuServerCreate(C.byref(h), C.byref(cfg)))
        self.handle = h.value
        self.rings = []

    def open_stream(self, nstation, nfrequency, blocks_per_dump, compute=XGPU_INT8, dump_slots=4,
                    dump_device=False):
        """Returns (stream id, dump ring of dump_slots dumps): a HostBuffer, or
        with dump_device a DeviceBuffer (the dumps stay in HBM)."""
        ns2 = nstation * (nstation + 1) // 2
        dump_bytes = nfrequency * ns2 * 4 * 2 * 8
        if dump_device:
            ring = DeviceBuffer(dump_bytes * dump_slots)
        else:
            ring = HostBuffer(dump_bytes * dump_slots,
                              np.int64 if compute == XGPU_INT8 else np.float64)
        cfg = _lib.XGPUStreamConfig(nstation, nfrequency, compute, blocks_per_dump,
                                    1 if dump_device else 0)
        sid = C.c_int()
        check(core().xgpuServerOpenStream(self.handle, C.byref(cfg), ring.ptr, dump_slots,
                                          C.byref(sid)))
        assert core().xgpuServerDumpBytes(self.handle, sid.value) == dump_bytes
        self.rings.append(ring)
        return sid.value, ring

    def submit(self, sid: int, block, ntime: int) -> None:
        """Host block (HostBuffer or host pointer)."""
        ptr = block.ptr if isinstance(block, HostBuffer) else int(block)
        check(core().xgpuServerSubmit(self.handle, sid, ptr, ntime))

    def submit_device(self, sid: int, block, ntime: int) -> None:
        """Block already in device memory (DeviceBuffer or device pointer)."""
        ptr = block.ptr if isinstance(block, DeviceBuffer) else int(block)
        check(core().xgpuServerSubmitDevice(self.handle, sid, ptr, ntime))

    def drain(self) -> None:
        check(core().xgpuServerDrain(self.handle))

    def blocks_done(self, sid):
        return int(core().xgpuServerBlocksDone(self.handle, sid))

    def dumps_done(self, sid):
        return int(core().xgpuServerDumpsDone(self.handle, sid))

    def stats(self) -> dict:
        st = _lib.XGPUServerStats()
        check(core().xgpuServerStats(self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.XGPUServerStats._fields_}

    def reset_stats(self) -> None:
        check(core().xgpuServerResetStats(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            core().xgpuServerDestroy(self.handle)
            self.handle = None
        for r in self.rings:
            r.free()
        self.rings = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
