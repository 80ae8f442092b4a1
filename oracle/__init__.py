"""CPU ORACLE of the X-engine -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
may import this package, and only as the checker (never as the measured or
shipped path).  It binds ``oracle/liboracle.so`` (the C restatement in
``xgpu_oracle.c``; how each function is pinned is listed in ``xgpu_oracle.h``)
and, when built, the reference's own CPU files compiled under ``oracle/_ref``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        P, I, LL = C.c_void_p, C.c_int, C.c_longlong
        for name, args in {
            "oracle_xengine_f32": [P, P, I, I, I, I],
            "oracle_xengine_dp4a_f32acc": [P, P, I, I, I, I],
            "oracle_xengine_i64": [P, P, I, I, I, I],
            "oracle_xengine_i64_pairs": [P, P, I, I, I, P, I, I],
            "oracle_xengine_i64_channels": [P, P, I, I, I, P, I, I],
            "oracle_xengine_f64_f16": [P, P, I, I, I, I],
            "oracle_reorder": [I, I, P, P, I, I, LL],
            "oracle_regtile_from_tri": [I, P, P, I, I, LL],
            "oracle_swizzle": [P, P, I, I, I],
            "oracle_unblock32": [P, P, C.c_size_t],
            "oracle_fft_f64": [P, P, I, I, I],
            "oracle_extract": [P, P, I, I],
            "oracle_random_legacy": [P, C.c_size_t, C.c_uint],
            "oracle_fill_uniform": [P, C.c_size_t, C.c_uint64, I, I],
        }.items():
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = args
        L.oracle_check.restype = I
        L.oracle_check.argtypes = [P, P, I, I, C.c_double, I, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def nbaseline(ns: int) -> int:
    return ns * (ns + 1) // 2


def tri_len(ns: int, nf: int) -> int:
    return nf * nbaseline(ns) * 4


def regtile_len(ns: int, nf: int) -> int:
    return nf * (ns // 2 + 1) * (ns // 4) * 16


def _check_input(x: np.ndarray, ns: int, nf: int, nt: int, dtype, per=2) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=dtype).reshape(-1)
    assert x.size == nt * nf * ns * 2 * per, (x.size, nt, nf, ns)
    return x


def xengine_f32(x, ns, nf, nt, nthreads=0) -> np.ndarray:
    """Legacy float X-engine (reference omp_xengine.c); triangular, interleaved."""
    x = _check_input(x, ns, nf, nt, np.int8)
    out = np.empty(2 * tri_len(ns, nf), dtype=np.float32)
    lib().oracle_xengine_f32(_p(out), _p(x), ns, nf, nt, nthreads)
    return out


def xengine_dp4a_f32acc(x, ns, nf, nt, nthreads=0) -> np.ndarray:
    x = _check_input(x, ns, nf, nt, np.int8)
    out = np.empty(2 * tri_len(ns, nf), dtype=np.int32)
    lib().oracle_xengine_dp4a_f32acc(_p(out), _p(x), ns, nf, nt, nthreads)
    return out


def xengine_i64(x, ns, nf, nt, nthreads=0) -> np.ndarray:
    """Exact X-engine; triangular order, interleaved int64 (re, im)."""
    x = _check_input(x, ns, nf, nt, np.int8)
    out = np.empty(2 * tri_len(ns, nf), dtype=np.int64)
    lib().oracle_xengine_i64(_p(out), _p(x), ns, nf, nt, nthreads)
    return out


def xengine_i64_pairs(x, ns, nf, nt, fij, nthreads=0) -> np.ndarray:
    """Exact visibilities of selected (f, i, j): shape (n, 4, 2) int64."""
    x = _check_input(x, ns, nf, nt, np.int8)
    fij = np.ascontiguousarray(fij, dtype=np.int32).reshape(-1, 3)
    out = np.empty((fij.shape[0], 4, 2), dtype=np.int64)
    lib().oracle_xengine_i64_pairs(_p(out), _p(x), ns, nf, nt, _p(fij), fij.shape[0], nthreads)
    return out


def xengine_i64_channels(x, ns, nf, nt, chans, nthreads=0) -> np.ndarray:
    """Exact visibilities of whole channels: (len(chans), ns(ns+1)/2, 4, 2)
    int64, each channel as xengine_i64 lays it out."""
    x = _check_input(x, ns, nf, nt, np.int8)
    ch = np.ascontiguousarray(chans, dtype=np.int32).reshape(-1)
    assert ((ch >= 0) & (ch < nf)).all()
    out = np.empty((ch.size, nbaseline(ns), 4, 2), dtype=np.int64)
    lib().oracle_xengine_i64_channels(_p(out), _p(x), ns, nf, nt, _p(ch), ch.size, nthreads)
    return out


def xengine_f64_f16(x16, ns, nf, nt, nthreads=0) -> np.ndarray:
    """FP64 X-engine of fp16 complex input (uint16 or float16 array)."""
    x = np.ascontiguousarray(x16).view(np.uint16).reshape(-1)
    assert x.size == nt * nf * ns * 4
    out = np.empty(2 * tri_len(ns, nf), dtype=np.float64)
    lib().oracle_xengine_f64_f16(_p(out), _p(x), ns, nf, nt, nthreads)
    return out


def reorder(order: int, src: np.ndarray, ns: int, nf: int) -> np.ndarray:
    """Split-plane REGISTER_TILE (3000) / REAL_IMAG (2000) -> interleaved triangular."""
    src = np.ascontiguousarray(src)
    plane = src.size // 2
    out = np.empty(2 * tri_len(ns, nf), dtype=src.dtype)
    lib().oracle_reorder(order, src.itemsize, _p(out), _p(src), ns, nf, plane)
    return out


def regtile_from_tri(tri: np.ndarray, ns: int, nf: int) -> np.ndarray:
    tri = np.ascontiguousarray(tri)
    plane = regtile_len(ns, nf)
    out = np.empty(2 * plane, dtype=tri.dtype)
    lib().oracle_regtile_from_tri(tri.itemsize, _p(out), _p(tri), ns, nf, plane)
    return out


def swizzle(x, ns, nf, nt) -> np.ndarray:
    x = _check_input(x, ns, nf, nt, np.int8)
    out = np.empty_like(x)
    lib().oracle_swizzle(_p(out), _p(x), ns, nf, nt)
    return out


def unblock32(x: np.ndarray) -> np.ndarray:
    """COMPLEX_BLOCK_SIZE=32 int8 input -> plain [t][f][s][p][re,im] input."""
    x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1)
    if x.size % 64:
        raise ValueError("block-32 input is a whole number of 64-byte blocks")
    out = np.empty_like(x)
    lib().oracle_unblock32(_p(out), _p(x), x.size)
    return out


def block32(x: np.ndarray) -> np.ndarray:
    """Inverse of unblock32: plain input -> COMPLEX_BLOCK_SIZE=32 layout."""
    x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1, 32, 2)
    return np.ascontiguousarray(x.transpose(0, 2, 1)).reshape(-1)


def fft_f64(x: np.ndarray, n: int, inverse: bool = False) -> np.ndarray:
    """Double-precision reference FFT (radix-2 Stockham) of complex64 data
    holding whole transforms of n points; returns complex128."""
    x = np.ascontiguousarray(x, dtype=np.complex64).reshape(-1)
    if x.size % n:
        raise ValueError("whole transforms only")
    out = np.empty(x.size, dtype=np.complex128)
    lib().oracle_fft_f64(_p(out), _p(x), n, x.size // n, int(inverse))
    return out


def fft_relative_error_ulps(reference: np.ndarray, result: np.ndarray, n: int) -> float:
    """Error measure of the reference FFT harness (src/fft/main.cpp:125-143):
    per transform sum|ref - res|^2 / sum|ref|^2, the maximum over the batch,
    square-rooted, in units of 2^-23."""
    ref = np.asarray(reference, dtype=np.complex128).reshape(-1, n)
    res = np.asarray(result, dtype=np.complex128).reshape(-1, n)
    err = (np.abs(ref - res) ** 2).sum(axis=1) / (np.abs(ref) ** 2).sum(axis=1)
    return float(np.sqrt(err.max()) / 1.192092896e-07)


def fft_harness_transpose(src: np.ndarray, sz: int, sy: int, sx: int) -> np.ndarray:
    """The reference FFT harness's record transpose (src/fft/main.cpp:145-165):
    per record of sx*sy*sz elements, out[sx*(z + y*sz) + x] = in[sx*(z*sy + y) + x]."""
    src = np.asarray(src).reshape(-1, sz, sy, sx)
    return np.ascontiguousarray(src.transpose(0, 2, 1, 3)).reshape(-1)


def fft_to_prototype(x: np.ndarray, n: int, spectrum: bool = False) -> np.ndarray:
    """Natural-order transforms -> the layout the reference FFT entry points
    read or write, as its harness arranges it (src/fft/main.cpp:249-264):
    n = 8, 16, 64 interleaved within groups; n >= 2048 spectra transposed
    (spectrum=True: the forward output / inverse input)."""
    x = np.asarray(x).reshape(-1)
    if n == 16:
        return fft_harness_transpose(x, 64, 16, 1)
    if n < 256:
        return fft_harness_transpose(x, 512 // n, 8, n // 8)
    if n >= 2048 and spectrum:
        return fft_harness_transpose(x, 512, n // 512, 1)
    return x.copy()


def fft_from_prototype(x: np.ndarray, n: int, spectrum: bool = False) -> np.ndarray:
    """Inverse of fft_to_prototype (the harness's transposes back)."""
    x = np.asarray(x).reshape(-1)
    if n == 16:
        return fft_harness_transpose(x, 16, 64, 1)
    if n < 256:
        return fft_harness_transpose(x, 8, 512 // n, n // 8)
    if n >= 2048 and spectrum:
        return fft_harness_transpose(x, n // 512, 512, 1)
    return x.copy()


def extract(tri: np.ndarray, ns: int, nf: int) -> np.ndarray:
    tri = np.ascontiguousarray(tri, dtype=np.float32)
    out = np.empty(nf * ns * ns * 4 * 2, dtype=np.float32)
    lib().oracle_extract(_p(out), _p(tri), ns, nf)
    return out


def random_legacy(ncomplex: int, seed: int) -> np.ndarray:
    out = np.empty(2 * ncomplex, dtype=np.int8)
    lib().oracle_random_legacy(_p(out), ncomplex, seed)
    return out


def fill_uniform(n: int, seed: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n, dtype=np.int8)
    lib().oracle_fill_uniform(_p(out), n, seed, lo, hi)
    return out


def check(gpu_tri: np.ndarray, cpu_tri: np.ndarray, ns: int, nf: int, tol=1e-5, round_gpu=True):
    """(error count, max relative error) as the reference xgpuCheckResult computes them."""
    g = np.ascontiguousarray(gpu_tri, dtype=np.float32)
    c = np.ascontiguousarray(cpu_tri, dtype=np.float32)
    mx = C.c_double()
    n = lib().oracle_check(_p(g), _p(c), ns, nf, tol, int(round_gpu), C.byref(mx))
    return n, mx.value


# ---------------------------------------------------------------------------
# The reference's own CPU files, compiled under oracle/_ref (may be absent).
# ---------------------------------------------------------------------------
def ref_path(tag: str) -> str:
    return os.path.join(HERE, "_ref", f"libxgpuref_{tag}.so")


def ref_lib(tag: str):
    """ctypes handle of oracle/_ref/libxgpuref_<tag>.so, or None if not built."""
    path = ref_path(tag)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    P = C.c_void_p
    L.xgpuOmpXengine.argtypes = [P, P]
    L.xgpuRandomComplex.argtypes = [P, C.c_ulonglong]
    L.xgpuReorderMatrix.argtypes = [P]
    L.xgpuSwizzleInput.argtypes = [P, P]
    L.xgpuExtractMatrix.argtypes = [P, P]
    L.xgpuCheckResult.argtypes = [P, P, C.c_int, P]
    for fn in ("xgpuOmpXengine", "xgpuRandomComplex", "xgpuReorderMatrix", "xgpuSwizzleInput",
               "xgpuExtractMatrix", "xgpuCheckResult"):
        getattr(L, fn).restype = None
    return L
