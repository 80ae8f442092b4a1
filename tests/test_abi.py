"""CPU tests of the C-ABI boundary: libraries load, every symbol declared in
include/*.h is exported, struct layouts match the reference header, sizes and
error codes behave as in the reference, and the CPU functions of the xgpu.h
API reproduce the reference's golden vectors.  No GPU compute calls."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import _lib
from conftest import ROOT, golden

INCLUDE = os.path.join(ROOT, "include")


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(xgpu\w+)\s*\(", text)
    return sorted(set(names))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_xgpu_h_declares_the_reference_api():
    names = declared_functions("xgpu.h")
    assert names == sorted([
        "xgpuVersionString", "xgpuInfo", "xgpuInit", "xgpuClearDeviceIntegrationBuffer",
        "xgpuSetHostInputBuffer", "xgpuSetHostOutputBuffer", "xgpuFree", "xgpuCudaXengine",
        "xgpuRandomComplex", "xgpuReorderMatrix", "xgpuCheckResult", "xgpuSwizzleInput",
        "xgpuExtractMatrix", "xgpuOmpXengine"])


@pytest.mark.parametrize("variant", ["libxgpu", "libxgpu_dp4a", "libxgpu_ns512_dp4a",
                                     "libxgpu_ns16", "libxgpu_ns16_ri", "libxgpu_ns16_tri",
                                     "libxgpu_ns16_cb32", "libxgpu_ns32_cb32_dp4a"])
def test_drop_in_library_exports_every_xgpu_h_symbol(variant):
    have = exported(_lib.lib_path(variant + ".so"))
    missing = [n for n in declared_functions("xgpu.h") if n not in have]
    assert not missing, missing
    X.variant(variant)  # loads, and its dependency on libxgpuamd.so resolves


# library that exports each extension header's functions
HEADER_LIBS = {"xgpu_amd.h": "libxgpuamd.so", "xgpu_serve.h": "libxgpuamd.so",
               "xgpu_fft.h": "libxgpufft.so"}


def test_core_library_exports_every_extension_symbol():
    headers = sorted(h for h in os.listdir(INCLUDE) if h != "xgpu.h")
    assert headers == sorted(HEADER_LIBS), headers
    for header in headers:
        have = exported(_lib.lib_path(HEADER_LIBS[header]))
        missing = [n for n in declared_functions(header) if n not in have]
        assert not missing, (header, missing)
    X.core()
    _lib.fft_lib()


def test_fft_library_declares_the_reference_entry_points():
    """The FFT prototype's entry points (src/fft/main.cpp:37-55)."""
    assert declared_functions("xgpu_fft.h") == ["xgpuAmdFft", "xgpuAmdFftTime"]
    sizes = (8, 16, 64, 256, 512, 1024, 2048, 4096, 8192)
    text = open(os.path.join(INCLUDE, "xgpu_fft.h")).read()
    for n in sizes:
        assert f"void FFT{n}(void *work, int batch);" in text
        assert f"void IFFT{n}(void *work, int batch);" in text
    have = exported(_lib.lib_path("libxgpufft.so"))
    assert all(f"FFT{n}" in have and f"IFFT{n}" in have for n in sizes)


def test_struct_layouts_match_reference_header():
    # reference src/xgpu.h:61-113 on LP64
    assert C.sizeof(_lib.XGPUInfo) == 88
    assert C.sizeof(_lib.XGPUContext) == 56
    assert _lib.XGPUContext.internal.offset == 48
    assert _lib.XGPUInfo.vecLength.offset == 32
    assert _lib.XGPUInfo.matrix_order.offset == 64


def sizing(variant):
    info = _lib.XGPUInfo()
    X.variant(variant).xgpuInfo(C.byref(info))
    return info


@pytest.mark.parametrize("variant,ns,nf,nt,ntp,order,compute", [
    ("libxgpu", 256, 10, 1024, 128, 3000, X.XGPU_FLOAT32),
    ("libxgpu_dp4a", 256, 10, 1024, 128, 3000, X.XGPU_INT8),
    ("libxgpu_ns512_dp4a", 512, 64, 8192, 256, 3000, X.XGPU_INT8),
    ("libxgpu_ns16_ri", 16, 2, 64, 16, 2000, X.XGPU_FLOAT32),
    ("libxgpu_ns16_tri", 16, 2, 64, 16, 1000, X.XGPU_FLOAT32),
])
def test_xgpuinfo_sizes(variant, ns, nf, nt, ntp, order, compute):
    """Same formulas as reference src/cuda_xengine.cu:113-142."""
    i = sizing(variant)
    assert (i.npol, i.nstation, i.nfrequency, i.ntime, i.ntimepipe) == (2, ns, nf, nt, ntp)
    assert i.nbaseline == (ns + 1) * (ns // 2)
    assert i.vecLength == nf * nt * ns * 2
    assert i.vecLengthPipe == nf * ntp * ns * 2
    assert i.triLength == nf * (ns + 1) * (ns // 2) * 4
    reg = nf * (ns // 2 + 1) * (ns // 4) * 16
    assert i.matLength == (reg if order == 3000 else i.triLength)
    assert i.matrix_order == order and i.input_type == X.XGPU_INT8
    assert i.compute_type == compute
    assert (i.shared_atomic_size, i.complex_block_size) == (4, 1)


@pytest.mark.parametrize("variant,ns,compute", [("libxgpu_ns16_cb32", 16, X.XGPU_FLOAT32),
                                                ("libxgpu_ns32_cb32_dp4a", 32, X.XGPU_INT8)])
def test_block32_builds_report_their_layout(variant, ns, compute):
    i = sizing(variant)
    assert (i.nstation, i.complex_block_size, i.compute_type) == (ns, 32, compute)
    r = X.info(ns, 2, compute=compute, complex_block_size=32)
    assert r.complex_block_size == 32 and X.info(ns, 2).complex_block_size == 1


def test_runtime_info_matches_build_time_info():
    for variant, ns, nf, nt, ntp, compute in (("libxgpu", 256, 10, 1024, 128, X.XGPU_FLOAT32),
                                              ("libxgpu_ns512_dp4a", 512, 64, 8192, 256, X.XGPU_INT8)):
        a = sizing(variant)
        b = X.info(ns, nf, nt, ntp, compute)
        for f, _ in _lib.XGPUInfo._fields_:
            assert getattr(a, f) == getattr(b, f), f


def test_runtime_info_rejects_bad_configs():
    for kw in (dict(nstation=7, nfrequency=1), dict(nstation=0, nfrequency=1),
               dict(nstation=6, nfrequency=1),          # register tile needs nstation % 4
               dict(nstation=16, nfrequency=0),
               dict(nstation=16, nfrequency=1, ntime=100, ntimepipe=33),
               dict(nstation=16, nfrequency=1, compute=9),
               dict(nstation=16, nfrequency=1, complex_block_size=16),
               dict(nstation=24, nfrequency=1, complex_block_size=32),  # needs nstation % 16
               dict(nstation=16, nfrequency=1, compute=X.XGPU_FLOAT16, complex_block_size=32)):
        with pytest.raises(X.XgpuError) as e:
            X.info(**kw)
        assert e.value.code == _lib.XGPU_INVALID_CONFIG
    X.info(6, 1, order=X.TRIANGULAR_ORDER)  # fine in triangular order


def test_xgpuinfo_tool_text_is_what_autoconf_parses():
    out = subprocess.run([_lib.lib_path("xgpuinfo")], capture_output=True, text=True,
                         check=True).stdout
    # the sed expressions of the reference's AX_CHECK_XGPUINFO (src/xgpu.m4:57-67)
    def sed(prefix):
        for line in out.splitlines():
            if prefix in line:
                return line.split(": ", 1)[1]
    assert sed("Number of stations: ") == "256"
    assert sed("Number of frequencies: ") == "10"
    assert sed("time samples per GPU integration: ") == "1024"
    assert "Output matrix order: register tile triangular" in out
    assert "Type of computation: FP32 multiply, FP32 accumulate" in out


def test_version_string():
    assert X.variant("libxgpu").xgpuVersionString().startswith(b"xgpu-mi355x")


def test_uninitialised_context_error_codes():
    lib = X.variant("libxgpu")
    ctx = _lib.XGPUContext()
    ctx.internal = None
    assert lib.xgpuClearDeviceIntegrationBuffer(C.byref(ctx)) == _lib.XGPU_NOT_INITIALIZED
    assert lib.xgpuSetHostInputBuffer(C.byref(ctx)) == _lib.XGPU_NOT_INITIALIZED
    assert lib.xgpuSetHostOutputBuffer(C.byref(ctx)) == _lib.XGPU_NOT_INITIALIZED
    assert lib.xgpuCudaXengine(C.byref(ctx), 1) == _lib.XGPU_NOT_INITIALIZED
    lib.xgpuFree(C.byref(ctx))  # no-op


def test_no_device_is_an_error_not_an_exit():
    """Reference src/cuda_xengine.cu:210-213 calls exit(-1); the drop-in returns.
    Run in a child process with no visible GPU."""
    code = ("import ctypes as C, sys; sys.path.insert(0, %r)\n"
            "import xgpu_amd as X\nfrom xgpu_amd import _lib\n"
            "ctx = _lib.XGPUContext()\n"
            "rc = X.variant('libxgpu').xgpuInit(C.byref(ctx), 0)\n"
            "print('rc', rc, ctx.internal)\n") % ROOT
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "rc 2 None" in r.stdout


def test_engine_without_device_reports_no_device():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    code = ("import sys; sys.path.insert(0, %r)\nimport xgpu_amd as X\n"
            "try:\n  X.Engine(16, 1)\nexcept X.XgpuError as e:\n  print('code', e.code)\n") % ROOT
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert "code 7" in r.stdout, r.stdout + r.stderr


def test_mfma_peak_data_rejects_bad_arguments():
    """Argument checks run before any device call (no GPU needed)."""
    lib = X.core()
    v = C.c_double()
    assert lib.xgpuAmdMfmaPeakData(0, None, 0, 3, C.byref(v)) == 6     # XGPU_INVALID_CONFIG
    assert lib.xgpuAmdMfmaPeakData(0, 4096, 0, 0, C.byref(v)) == 6
    assert lib.xgpuAmdMfmaPeakData(0, 4096, 0, 3, None) == 6
    assert lib.xgpuAmdMfmaPeakData(0, 4096, 1, 3, C.byref(v)) == 6     # FP32: no MFMA probe


# ---------------------------------------------------------------- CPU functions of xgpu.h
libc = C.CDLL("libc.so.6")


def test_random_complex_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    lib = X.variant("libxgpu_ns16")
    n = 2 * 64 * 16 * 2
    out = np.zeros(2 * n, dtype=np.int8)
    libc.srand(1)
    lib.xgpuRandomComplex(out.ctypes.data, n)
    assert np.array_equal(out, g["x_legacy"])


def test_omp_xengine_float_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    lib = X.variant("libxgpu_ns16")
    for key_in, key_out in (("x_legacy", "omp_legacy"), ("x_full", "omp_full")):
        out = np.zeros_like(g[key_out])
        x = np.ascontiguousarray(g[key_in])
        lib.xgpuOmpXengine(out.ctypes.data, x.ctypes.data)
        assert np.array_equal(out, g[key_out])


@pytest.mark.parametrize("variant,tag,dtype", [
    ("libxgpu_ns16_cb32", "ns16_nf2_nt64_cb32", np.float32),
    ("libxgpu_ns32_cb32_dp4a", "ns32_nf3_nt256_cb32_dp4a", np.int32)])
def test_omp_xengine_block32_matches_reference(meta, variant, tag, dtype):
    """COMPLEX_BLOCK_SIZE=32 builds read the blocked input like the reference
    built the same way (golden output of its xgpuOmpXengine)."""
    m = meta[tag]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    ref = golden(f"ref_{tag}.npz")["omp"]
    out = np.zeros_like(ref)
    X.variant(variant).xgpuOmpXengine(out.ctypes.data, x.ctypes.data)
    assert out.dtype == dtype and np.array_equal(out, ref)


def test_omp_xengine_dp4a_is_exact():
    """DP4A build: exact integer sums (the reference rounds through float past
    2^24; tests/test_oracle.py pins that behaviour separately)."""
    ns, nf, nt = 16, 2, 8192
    x = oracle.fill_uniform(nt * nf * ns * 4, 3, -128, 127)
    out = np.zeros(2 * oracle.tri_len(ns, nf), dtype=np.int32)
    X.variant("libxgpu_ns16_dp4a").xgpuOmpXengine(out.ctypes.data, x.ctypes.data)
    assert np.array_equal(out.astype(np.int64), oracle.xengine_i64(x, ns, nf, nt))


def test_reorder_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    m = g["regtile"].copy()
    X.variant("libxgpu_ns16").xgpuReorderMatrix(m.ctypes.data)
    n = 2 * oracle.tri_len(16, 2)
    assert np.array_equal(m[:n], g["reordered"][:n])
    assert not m[n:].any()  # the tail is zero here (unspecified in the reference)
    g = golden("ref_ns16_nf2_nt64_ri.npz")
    m = g["realimag"].copy()
    X.variant("libxgpu_ns16_ri").xgpuReorderMatrix(m.ctypes.data)
    assert np.array_equal(m, g["reordered"])
    m = np.arange(2 * oracle.tri_len(16, 2), dtype=np.float32)
    m2 = m.copy()
    X.variant("libxgpu_ns16_tri").xgpuReorderMatrix(m2.ctypes.data)
    assert np.array_equal(m, m2)  # triangular order: nothing to do


def test_extract_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    full = np.zeros_like(g["full"])
    packed = np.ascontiguousarray(g["omp_legacy"])
    X.variant("libxgpu_ns16").xgpuExtractMatrix(full.ctypes.data, packed.ctypes.data)
    assert np.array_equal(full, g["full"])


def capture_fd1(fn):
    import sys
    import tempfile
    libc.fflush(None)
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            fn()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def test_swizzle_matches_reference(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    out = np.zeros_like(g["x_legacy"])
    x = np.ascontiguousarray(g["x_legacy"])
    text = capture_fd1(lambda: X.variant("libxgpu_ns16").xgpuSwizzleInput(out.ctypes.data, x.ctypes.data))
    assert np.array_equal(out, g["swizzled"])
    assert text == meta["ns16_nf2_nt64"]["swizzle_text"]


def test_check_result_report_matches_reference(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    lib = X.variant("libxgpu_ns16")
    cpu = np.ascontiguousarray(g["omp_legacy"])
    for verbose in ("0", "1"):
        gpu = g["check_gpu"].copy()
        text = capture_fd1(lambda: lib.xgpuCheckResult(gpu.ctypes.data, cpu.ctypes.data,
                                                       int(verbose), None))
        assert text == meta["ns16_nf2_nt64"]["check_text"][verbose]
    gpu = cpu.copy()
    text = capture_fd1(lambda: lib.xgpuCheckResult(gpu.ctypes.data, cpu.ctypes.data, 0, None))
    assert text == meta["ns16_nf2_nt64"]["check_text"]["ok"]
