"""GPU parity of the batched FFTs (libxgpufft.so, include/xgpu_fft.h).

Against the double-precision reference FFT (oracle.fft_f64, the reference
harness's own reference, src/fft/main.cpp:80-117), with the harness's error
measure (src/fft/main.cpp:125-143: per-transform relative RMS error in units
of 2^-23, the worst transform of the batch).  Bounds:
  * Higham's worst-case bound for a radix-2-equivalent FFT in floating point
    (Accuracy and Stability of Numerical Algorithms, Th. 24.2; cited by the
    reference harness) is about log2(N) (mu + gamma_4 (sqrt(2) + mu)) with
    twiddle error mu <= u = 2^-24, i.e. below 3.4 log2(N) units of 2^-23;
  * the reference expects errors "of order 1"; the tests require <= 2.5.
The prototype layouts of the reference entry points are checked through the
harness's own transposes (oracle.fft_to_prototype / fft_from_prototype), and
the two-way test of the harness (FFT then IFFT on untransposed data, divided
by N) is run for every size.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import xgpu_amd as X
from xgpu_amd import _lib
from conftest import flush

pytestmark = pytest.mark.gpu

ULP_TARGET = 2.5
PROTO_SIZES = (8, 16, 64, 256, 512, 1024, 2048, 4096, 8192)


def rand_c64(n, seed):
    g = np.random.default_rng(seed)
    return (g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)).astype(np.complex64)


def check_ulps(ref, got, n):
    ulps = oracle.fft_relative_error_ulps(ref, got, n)
    assert ulps <= min(ULP_TARGET, 3.4 * math.log2(n)), ulps
    return ulps


def run_natural(x, n, inverse):
    buf = X.DeviceBuffer.from_array(x)
    X.fft(buf, n, x.size // n, inverse=inverse)
    got = buf.download(np.complex64)
    buf.free()
    return got


@pytest.mark.parametrize("n", X.FFT_SIZES)
def test_fft_natural_forward_inverse_and_round_trip(n):
    # a batch that leaves the last workgroup partly filled (2048 / n transforms per workgroup)
    per_wg = max(1, 2048 // n)
    batch = max(3, (1 << 18) // n) + (per_wg // 2 if per_wg > 1 else 0)
    x = rand_c64(n * batch, 100 + n)
    flush(f"[fft] natural n={n} batch={batch}")
    fwd = run_natural(x, n, False)
    check_ulps(oracle.fft_f64(x, n), fwd, n)
    inv = run_natural(x, n, True)
    check_ulps(oracle.fft_f64(x, n, inverse=True), inv, n)
    back = run_natural(fwd, n, True) / n
    check_ulps(x, back, n)


def test_fft_closed_forms_on_device():
    n, batch = 4096, 3
    x = np.zeros((batch, n), np.complex64)
    x[0, 0] = 1                                                    # impulse -> ones
    x[1, :] = 1                                                    # constant -> n at bin 0
    x[2, :] = np.exp(2j * np.pi * 5 * np.arange(n) / n)            # tone -> bin 5
    got = run_natural(x.reshape(-1), n, False).reshape(batch, n)
    assert np.abs(got[0] - 1).max() < 1e-6
    assert abs(got[1, 0] - n) < 1e-3 and np.abs(got[1, 1:]).max() < 1e-3
    assert abs(got[2, 5] - n) < 1e-2 and np.abs(np.delete(got[2], 5)).max() < 1e-2


@pytest.mark.parametrize("n", PROTO_SIZES)
def test_reference_entry_points_in_prototype_layouts(n):
    """FFT<N> / IFFT<N> read and write the layouts the reference harness
    transposes to and from; results match the double-precision reference."""
    lib = _lib.fft_lib()
    group = 64 if n in (8, 16) else 8 if n == 64 else 1
    batch = max(group, (1 << 17) // n)
    x = rand_c64(n * batch, 200 + n)
    flush(f"[fft] prototype n={n} batch={batch}")
    ref = oracle.fft_f64(x, n)
    buf = X.DeviceBuffer.from_array(oracle.fft_to_prototype(x, n))
    getattr(lib, f"FFT{n}")(buf.ptr, batch)
    X.core().xgpuAmdDeviceSynchronize()
    spec = oracle.fft_from_prototype(buf.download(np.complex64), n, spectrum=True)
    check_ulps(ref, spec, n)
    # the inverse reads the forward output's layout and writes the input's
    getattr(lib, f"IFFT{n}")(buf.ptr, batch)
    X.core().xgpuAmdDeviceSynchronize()
    back = oracle.fft_from_prototype(buf.download(np.complex64), n) / n
    check_ulps(x, back, n)
    buf.free()


@pytest.mark.parametrize("n", PROTO_SIZES)
def test_harness_two_way_on_untransposed_data(n):
    """The harness's two-way test (src/fft/main.cpp:297-318): FFT then IFFT on
    the untransposed source, divided by N, returns the source."""
    lib = _lib.fft_lib()
    group = 64 if n in (8, 16) else 8 if n == 64 else 1
    batch = max(group, (1 << 16) // n)
    x = rand_c64(n * batch, 300 + n)
    buf = X.DeviceBuffer.from_array(x)
    getattr(lib, f"FFT{n}")(buf.ptr, batch)
    getattr(lib, f"IFFT{n}")(buf.ptr, batch)
    X.core().xgpuAmdDeviceSynchronize()
    back = buf.download(np.complex64) / n
    buf.free()
    check_ulps(x, back, n)


def test_rejected_calls_leave_data_untouched():
    lib = _lib.fft_lib()
    x = rand_c64(8 * 100, 9)
    buf = X.DeviceBuffer.from_array(x)
    assert lib.xgpuAmdFft(buf.ptr, 8, 100, 0, X.FFT_PROTOTYPE, None) == 2   # not a multiple of 64
    assert lib.xgpuAmdFft(buf.ptr, 24, 10, 0, X.FFT_NATURAL, None) == 1     # not a power of two
    assert lib.xgpuAmdFft(buf.ptr, 32, 10, 0, X.FFT_PROTOTYPE, None) == 1   # no prototype entry
    assert lib.xgpuAmdFft(buf.ptr, 8, 10, 0, 7, None) == 3                  # unknown layout
    assert lib.xgpuAmdFft(buf.ptr, 8, -1, 0, X.FFT_NATURAL, None) == 3
    assert lib.xgpuAmdFft(buf.ptr + 4, 8, 1, 0, X.FFT_NATURAL, None) == 3   # unaligned
    assert lib.xgpuAmdFft(buf.ptr, 8, 0, 0, X.FFT_NATURAL, None) == 0       # empty batch
    lib.FFT8(buf.ptr, 100)                                                   # prints, no-op
    X.core().xgpuAmdDeviceSynchronize()
    assert np.array_equal(buf.download(np.complex64), x)
    buf.free()
