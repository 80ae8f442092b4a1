"""ctypes bindings of the C-ABI libraries built in ``xgpu_amd/lib``.

``libxgpuamd.so`` carries the gfx950 kernels, the thin HIP layer and the
runtime-sized host runtime (``include/xgpu_amd.h``); ``libxgpu*.so`` are the
build-time-sized ``include/xgpu.h`` drop-in libraries.  Nothing here falls back
to a CPU implementation: a missing library raises ``XgpuLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

# xgpu.h constants (reference src/xgpu.h:40-127)
XGPU_INT8, XGPU_FLOAT32, XGPU_INT32, XGPU_FLOAT16 = 0, 1, 2, 3
TRIANGULAR_ORDER = 1000
REAL_IMAG_TRIANGULAR_ORDER = 2000
REGISTER_TILE_TRIANGULAR_ORDER = 3000
XGPU_DEVICE_MASK = (1 << 16) - 1
XGPU_DONT_REGISTER_ARRAY = 1 << 16
XGPU_DONT_REGISTER_MATRIX = 1 << 17
XGPU_DONT_REGISTER = XGPU_DONT_REGISTER_ARRAY | XGPU_DONT_REGISTER_MATRIX
XGPU_OK, XGPU_OUT_OF_MEMORY, XGPU_CUDA_ERROR = 0, 1, 2
XGPU_INSUFFICIENT_TEXTURE_MEMORY, XGPU_NOT_INITIALIZED, XGPU_HOST_BUFFER_NOT_SET = 3, 4, 5
XGPU_INVALID_CONFIG, XGPU_NO_DEVICE = 6, 7
SYNCOP_NONE, SYNCOP_DUMP, SYNCOP_SYNC_TRANSFER, SYNCOP_SYNC_COMPUTE = 0, 1, 2, 3
XGPU_DUMP_DEVICE, XGPU_DUMP_WIDE, XGPU_DUMP_TRIANGULAR = 1, 2, 4
XGPU_H2D, XGPU_D2H, XGPU_D2D = 1, 2, 3


class XgpuLibraryMissing(RuntimeError):
    """The native library has not been built (run ``__graft_entry__.build()``)."""


class XgpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"xgpu error {code}: {msg}")
        self.code = code


class XGPUInfo(C.Structure):
    # reference src/xgpu.h:61-92
    _fields_ = [
        ("npol", C.c_uint), ("nstation", C.c_uint), ("nbaseline", C.c_uint),
        ("nfrequency", C.c_uint), ("ntime", C.c_uint), ("ntimepipe", C.c_uint),
        ("input_type", C.c_uint), ("compute_type", C.c_uint),
        ("vecLength", C.c_ulonglong), ("vecLengthPipe", C.c_ulonglong),
        ("matLength", C.c_ulonglong), ("triLength", C.c_ulonglong),
        ("matrix_order", C.c_uint), ("shared_atomic_size", C.c_size_t),
        ("complex_block_size", C.c_size_t),
    ]


class XGPUContext(C.Structure):
    # reference src/xgpu.h:94-113
    _fields_ = [
        ("array_h", C.c_void_p), ("matrix_h", C.c_void_p),
        ("array_len", C.c_size_t), ("matrix_len", C.c_size_t),
        ("input_offset", C.c_size_t), ("output_offset", C.c_size_t),
        ("internal", C.c_void_p),
    ]


class XGPUAmdConfig(C.Structure):
    _fields_ = [
        ("npol", C.c_uint), ("nstation", C.c_uint), ("nfrequency", C.c_uint),
        ("ntime", C.c_uint), ("ntimepipe", C.c_uint), ("compute_type", C.c_uint),
        ("matrix_order", C.c_uint), ("device", C.c_int), ("complex_block_size", C.c_uint),
        ("time_swizzle", C.c_uint),
    ]


class XGPUAmdBatchItem(C.Structure):
    _fields_ = [("engine", C.c_void_p), ("input_d", C.c_void_p), ("ntime", C.c_uint)]


class XGPUAmdTuning(C.Structure):
    _fields_ = [("wg_per_cu", C.c_int), ("prefetch", C.c_int), ("graph", C.c_int),
                ("kernel", C.c_int), ("stage", C.c_int), ("staging", C.c_int),
                ("ntimepipe", C.c_int), ("mfma", C.c_int)]


class XGPUServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("staging_bytes", C.c_size_t), ("max_batch", C.c_int),
                ("max_inflight", C.c_int)]


class XGPUStreamConfig(C.Structure):
    _fields_ = [("nstation", C.c_uint), ("nfrequency", C.c_uint), ("compute_type", C.c_uint),
                ("blocks_per_dump", C.c_uint), ("dump_device", C.c_uint)]


class XGPUServerStats(C.Structure):
    _fields_ = [("blocks_submitted", C.c_ulonglong), ("blocks_done", C.c_ulonglong),
                ("dumps_done", C.c_ulonglong), ("launches", C.c_ulonglong),
                ("cmac_done", C.c_ulonglong), ("busy_s", C.c_double),
                ("tcmac_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_batch", C.c_double),
                ("mean_fused", C.c_double), ("staging_peak_bytes", C.c_size_t)]


_core = None
_variants: dict = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def _load(name: str) -> C.CDLL:
    path = lib_path(name)
    if not os.path.exists(path):
        raise XgpuLibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    return C.CDLL(path)


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


def core() -> C.CDLL:
    """The runtime-sized library (include/xgpu_amd.h)."""
    global _core
    if _core is None:
        lib = _load("libxgpuamd.so")
        P, I, U, SZ, ULL = C.c_void_p, C.c_int, C.c_uint, C.c_size_t, C.c_ulonglong
        _sig(lib, "xgpuAmdInfo", I, C.POINTER(XGPUAmdConfig), C.POINTER(XGPUInfo))
        _sig(lib, "xgpuAmdCreate", I, C.POINTER(P), C.POINTER(XGPUAmdConfig))
        _sig(lib, "xgpuAmdDestroy", None, P)
        _sig(lib, "xgpuAmdClear", I, P)
        _sig(lib, "xgpuAmdCorrelate", I, P, P, U)
        _sig(lib, "xgpuAmdCorrelateNoAlloc", I, P, P, U)
        _sig(lib, "xgpuAmdCorrelateHost", I, P, P, I, P)
        _sig(lib, "xgpuAmdDump", I, P, P, I)
        _sig(lib, "xgpuAmdSynchronize", I, P)
        _sig(lib, "xgpuAmdStream", P, P)
        _sig(lib, "xgpuAmdSamples", ULL, P)
        _sig(lib, "xgpuAmdLastError", C.c_char_p)
        _sig(lib, "xgpuAmdCorrelateBatch", I, C.POINTER(XGPUAmdBatchItem), I, P)
        _sig(lib, "xgpuAmdMalloc", I, C.POINTER(P), SZ)
        _sig(lib, "xgpuAmdFree", I, P)
        _sig(lib, "xgpuAmdMemcpy", I, P, P, SZ, I)
        _sig(lib, "xgpuAmdDeviceSynchronize", I)
        _sig(lib, "xgpuAmdFillUniformInt8", I, P, SZ, ULL, I, I)
        _sig(lib, "xgpuAmdSwizzleInput", I, P, P, U, U, U, I, P)
        _sig(lib, "xgpuAmdExtractMatrix", I, P, P, U, U, I, P)
        _sig(lib, "xgpuAmdTimeCorrelate", I, P, P, U, I, I, C.POINTER(C.c_float))
        _sig(lib, "xgpuAmdMfmaPeak", I, I, I, C.POINTER(C.c_double))
        _sig(lib, "xgpuAmdMfmaPeakData", I, I, P, U, I, C.POINTER(C.c_double))
        _sig(lib, "xgpuAmdSetTuning", I, P, C.POINTER(XGPUAmdTuning))
        _sig(lib, "xgpuAmdGetTuning", I, P, C.POINTER(XGPUAmdTuning))
        _sig(lib, "xgpuAmdAutotune", I, P, P, U, I, I, C.c_char_p, C.POINTER(XGPUAmdTuning),
             C.POINTER(C.c_float))
        _sig(lib, "xgpuAmdLoadTuning", I, P, C.c_char_p, U)
        _sig(lib, "xgpuAmdAutotuneHost", I, P, P, I, I, C.c_char_p, C.POINTER(XGPUAmdTuning),
             C.POINTER(C.c_float))
        _sig(lib, "xgpuAmdHostAlloc", I, C.POINTER(P), SZ)
        _sig(lib, "xgpuAmdHostFree", I, P)
        _sig(lib, "xgpuServerCreate", I, C.POINTER(P), C.POINTER(XGPUServerConfig))
        _sig(lib, "xgpuServerDestroy", I, P)
        _sig(lib, "xgpuServerOpenStream", I, P, C.POINTER(XGPUStreamConfig), P, U, C.POINTER(I))
        _sig(lib, "xgpuServerDumpBytes", SZ, P, I)
        _sig(lib, "xgpuServerSubmit", I, P, I, P, U)
        _sig(lib, "xgpuServerSubmitDevice", I, P, I, P, U)
        _sig(lib, "xgpuServerDrain", I, P)
        _sig(lib, "xgpuServerBlocksDone", ULL, P, I)
        _sig(lib, "xgpuServerDumpsDone", ULL, P, I)
        _sig(lib, "xgpuServerStats", I, P, C.POINTER(XGPUServerStats))
        _sig(lib, "xgpuServerResetStats", I, P)
        _core = lib
    return _core


_fft = None


def fft_lib() -> C.CDLL:
    """The batched FFT library (include/xgpu_fft.h)."""
    global _fft
    if _fft is None:
        lib = _load("libxgpufft.so")
        P, I = C.c_void_p, C.c_int
        _sig(lib, "xgpuAmdFft", I, P, I, C.c_longlong, I, I, P)
        _sig(lib, "xgpuAmdFftTime", I, P, I, C.c_longlong, I, I, I, C.POINTER(C.c_float))
        for n in (8, 16, 64, 256, 512, 1024, 2048, 4096, 8192):
            _sig(lib, f"FFT{n}", None, P, I)
            _sig(lib, f"IFFT{n}", None, P, I)
        _fft = lib
    return _fft


def variant(name: str = "libxgpu") -> C.CDLL:
    """A build-time-sized xgpu.h library, e.g. ``libxgpu`` or ``libxgpu_ns512_dp4a``."""
    if name not in _variants:
        core()  # resolve the dependency from this directory first
        lib = _load(name + ".so")
        P, I = C.c_void_p, C.c_int
        _sig(lib, "xgpuVersionString", C.c_char_p)
        _sig(lib, "xgpuInfo", None, C.POINTER(XGPUInfo))
        _sig(lib, "xgpuInit", I, C.POINTER(XGPUContext), I)
        _sig(lib, "xgpuClearDeviceIntegrationBuffer", I, C.POINTER(XGPUContext))
        _sig(lib, "xgpuSetHostInputBuffer", I, C.POINTER(XGPUContext))
        _sig(lib, "xgpuSetHostOutputBuffer", I, C.POINTER(XGPUContext))
        _sig(lib, "xgpuFree", None, C.POINTER(XGPUContext))
        _sig(lib, "xgpuCudaXengine", I, C.POINTER(XGPUContext), I)
        _sig(lib, "xgpuRandomComplex", None, P, C.c_ulonglong)
        _sig(lib, "xgpuReorderMatrix", None, P)
        _sig(lib, "xgpuCheckResult", None, P, P, I, P)
        _sig(lib, "xgpuSwizzleInput", None, P, P)
        _sig(lib, "xgpuExtractMatrix", None, P, P)
        _sig(lib, "xgpuOmpXengine", None, P, P)
        _variants[name] = lib
    return _variants[name]


def check(rc: int) -> None:
    if rc != XGPU_OK:
        msg = core().xgpuAmdLastError()
        raise XgpuError(rc, msg.decode() if msg else "")
