"""Pin the FFT oracle (oracle_fft_f64, the reference FFT harness's
double-precision Stockham FFT, src/fft/main.cpp:80-117) and the harness
layouts before trusting them.

Checks: numpy.fft (a library FFT) on random data, closed forms (impulse,
constant, single tone), Parseval, the forward/inverse round trip, the error
measure of src/fft/main.cpp:125-143 on planted errors, and the prototype
layouts against the harness's record transpose (src/fft/main.cpp:145-165,
249-264) restated independently.
"""
import numpy as np
import pytest

import oracle

SIZES = (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192)


def rand_c64(n, seed):
    g = np.random.default_rng(seed)
    return (g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)).astype(np.complex64)


@pytest.mark.parametrize("n", SIZES)
def test_fft_f64_matches_numpy(n):
    batch = max(1, 4096 // n)
    x = rand_c64(n * batch, n)
    want = np.fft.fft(x.astype(np.complex128).reshape(batch, n), axis=1).reshape(-1)
    assert np.allclose(oracle.fft_f64(x, n), want, rtol=0, atol=1e-11 * np.sqrt(n))
    winv = np.fft.ifft(x.astype(np.complex128).reshape(batch, n), axis=1).reshape(-1) * n
    assert np.allclose(oracle.fft_f64(x, n, inverse=True), winv, rtol=0, atol=1e-11 * np.sqrt(n))


def test_fft_f64_closed_forms():
    n = 1024
    imp = np.zeros(n, np.complex64)
    imp[0] = 1
    assert np.allclose(oracle.fft_f64(imp, n), 1.0)          # impulse -> constant
    const = np.ones(n, np.complex64)
    want = np.zeros(n)
    want[0] = n
    assert np.allclose(oracle.fft_f64(const, n), want)       # constant -> n at bin 0
    k0 = 37
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n).astype(np.complex64)
    spec = oracle.fft_f64(tone, n)
    assert abs(spec[k0] - n) < 1e-3 and np.abs(np.delete(spec, k0)).max() < 1e-3  # tone -> bin k0
    shifted = np.zeros(n, np.complex64)
    shifted[1] = 1                                            # delay by one: exp(-2 pi i k / n)
    assert np.allclose(oracle.fft_f64(shifted, n), np.exp(-2j * np.pi * np.arange(n) / n))


def test_fft_f64_parseval_and_round_trip():
    n, batch = 512, 8
    x = rand_c64(n * batch, 3)
    X = oracle.fft_f64(x, n)
    assert np.allclose((np.abs(X.reshape(batch, n)) ** 2).sum(1),
                       n * (np.abs(x.astype(np.complex128).reshape(batch, n)) ** 2).sum(1))
    back = oracle.fft_f64(X.astype(np.complex64), n, inverse=True) / n
    assert np.abs(back - x).max() < 1e-5


def test_relative_error_measure():
    n = 64
    ref = rand_c64(n * 3, 5).astype(np.complex128)
    assert oracle.fft_relative_error_ulps(ref, ref, n) == 0.0
    res = ref.copy()
    res[n + 5] += 1e-6 * np.sqrt((np.abs(ref[n:2 * n]) ** 2).sum())  # transform 1 only
    ulps = oracle.fft_relative_error_ulps(ref, res, n)
    assert abs(ulps - 1e-6 / 1.192092896e-07) < 1e-3   # the worst transform, in ULPs


def harness_transpose(src, sz, sy, sx):
    """Element-by-element restatement of the harness's record transpose."""
    src = np.asarray(src).reshape(-1)
    rec = sx * sy * sz
    out = np.empty_like(src)
    for r0 in range(0, src.size, rec):
        for z in range(sz):
            for y in range(sy):
                for x in range(sx):
                    out[r0 + sx * (z + y * sz) + x] = src[r0 + sx * (z * sy + y) + x]
    return out


@pytest.mark.parametrize("n", (8, 16, 64, 256, 512, 1024, 2048, 4096, 8192))
def test_prototype_layouts_are_the_harness_transposes(n):
    total = max(n, 2048)
    x = np.arange(total)
    if n == 16:
        to, back = (64, 16, 1), (16, 64, 1)
    elif n < 256:
        to, back = (512 // n, 8, n // 8), (8, 512 // n, n // 8)
    else:
        to, back = None, (n // 512, 512, 1) if n >= 2048 else None
    want_in = harness_transpose(x, *to) if to else x
    assert np.array_equal(oracle.fft_to_prototype(x, n), want_in)
    if n < 256:
        assert np.array_equal(oracle.fft_from_prototype(want_in, n), x)
    if n >= 2048:
        # the harness transposes the forward output back with `back`
        proto_spec = oracle.fft_to_prototype(x, n, spectrum=True)
        assert np.array_equal(harness_transpose(proto_spec, *back), x)
        # bin y (n/512) + z sits at offset 512 z + y
        z, y = 3, 17
        assert proto_spec[512 * z + y] == y * (n // 512) + z
