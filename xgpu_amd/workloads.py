"""Seeded synthetic inputs of the benchmark configurations, and the error
statistics the floating-point paths are reported with.

Shared by bench.py and the GPU tests so that both measure and check the same
data.  Input order is the reference's [t][f][s][p][re, im]
(reference src/omp_xengine.c:57-60).
"""
from __future__ import annotations

import numpy as np

SERVE_NS = (64, 96, 128, 200, 256, 352, 512)
SERVE_NT = (256, 512, 1024, 2048, 4096)
I32_SAMPLE_BOUND = 65535


def cfg3_input(ns=2048, nf=16, nt=4096, seed=3, sigma=20.0, amp=8.0, omega=0.0123) -> np.ndarray:
    """round(N(0, sigma^2)) per component plus a common tone amp e^{i(omega t +
    phi_s)} with a uniform random phase per station, clipped to [-127, 127]:
    int8 array (nt, nf, ns, 2, 2).  Generated channel by channel to bound host
    memory."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi, ns)
    x = np.empty((nt, nf, ns, 2, 2), dtype=np.int8)
    t = np.arange(nt, dtype=np.float32)[:, None, None]
    for f in range(nf):
        tone = amp * np.exp(1j * (omega * t + phi[None, None, :]))
        g = rng.normal(0, sigma, (nt, ns, 2, 2)).astype(np.float32)
        g[..., 0] += tone.real[:, 0, :, None].astype(np.float32)
        g[..., 1] += tone.imag[:, 0, :, None].astype(np.float32)
        x[:, f] = np.clip(np.rint(g), -127, 127).astype(np.int8)
    return x


def cfg4_input(ns=256, nf=128, nt=2048, seed=4, sigma=16.0, amp=4.0, omega=0.0371,
               dtype=np.float16) -> np.ndarray:
    """Complex N(0, sigma^2) plus a tone amp e^{i(2 pi f tau_s / NF + omega t)}
    with a random delay tau_s per station: array (nt, nf, ns, 2, 2) of `dtype`
    (float16 for the FP16 path).  With dtype=np.int8 the values are rounded and
    clipped to [-127, 127] instead: int8 data that the FP16 path and the FP32
    plain path both represent exactly, for comparing the two."""
    rng = np.random.default_rng(seed)
    tau = rng.uniform(0, 8, ns)
    t = np.arange(nt)[:, None, None, None]
    f = np.arange(nf)[None, :, None, None]
    ph = 2 * np.pi * f * tau[None, None, :, None] / nf + omega * t
    x = np.empty((nt, nf, ns, 2, 2), dtype=dtype)
    for t0 in range(0, nt, 256):  # in slabs to bound host memory
        sl = slice(t0, t0 + 256)
        n = min(256, nt - t0)
        g = rng.normal(0, sigma, (n, nf, ns, 2, 2))
        g[..., 0] += amp * np.cos(ph[sl])
        g[..., 1] += amp * np.sin(ph[sl])
        if dtype == np.int8:
            x[sl] = np.clip(np.rint(g), -127, 127).astype(np.int8)
        else:
            x[sl] = g.astype(dtype)
    return x


def serving_mix(seed=5, nstreams=8, spill_stream=True):
    """The serving mix: nstreams subarray streams with NSTATION from SERVE_NS,
    NFREQUENCY uniform in [8, 128], block NTIME from SERVE_NT and K uniform in
    [4, 64] blocks per dump, as (nstation, nfrequency, ntime, K) tuples.  With
    spill_stream, if no stream's integration (K * NTIME samples) passes the
    int32 bound, the stream with the longest blocks gets at least 2048-sample
    blocks and K large enough to pass it (its engine then spills to int64
    inside the integration)."""
    rng = np.random.default_rng(seed)
    streams = []
    for _ in range(nstreams):
        ns = int(rng.choice(SERVE_NS))
        nf = int(rng.integers(8, 129))
        nt = int(rng.choice(SERVE_NT))
        k = int(rng.integers(4, 65))
        streams.append((ns, nf, nt, k))
    if spill_stream and not any(nt * k > I32_SAMPLE_BOUND for _, _, nt, k in streams):
        # the longest-block stream integrates past the bound: 4096-sample
        # blocks need K >= 17, 2048-sample blocks K >= 33
        i = int(np.argmax([nt for _, _, nt, _ in streams]))
        ns, nf, nt, k = streams[i]
        nt = max(nt, 2048)
        streams[i] = (ns, nf, nt, max(k, I32_SAMPLE_BOUND // nt + 1))
    return streams


def vis_error_stats(got: np.ndarray, ref: np.ndarray, ns: int) -> dict:
    """Errors of visibilities against a reference, both interleaved triangular
    (re, im) arrays of whole channels (reshape (nchan, ns(ns+1)/2, 4, 2)):
      max_abs_err   max |dV| over every complex visibility,
      rms_rel_err   sqrt(sum |dV|^2 / sum |V|^2),
      max_norm_err  max |dV_ij| / sqrt(V_ii V_jj) per polarization pair (the
                    autocorrelations V_ii taken from `ref`)."""
    g = np.asarray(got, dtype=np.float64).reshape(-1, ns * (ns + 1) // 2, 4, 2)
    r = np.asarray(ref, dtype=np.float64).reshape(-1, ns * (ns + 1) // 2, 4, 2)
    dv = np.hypot(g[..., 0] - r[..., 0], g[..., 1] - r[..., 1])
    mag = np.hypot(r[..., 0], r[..., 1])
    i = np.repeat(np.arange(ns), np.arange(1, ns + 1))
    j = np.concatenate([np.arange(k + 1) for k in range(ns)])
    di = i * (i + 1) // 2 + i
    dj = j * (j + 1) // 2 + j
    auto = r[:, :, [0, 3], 0]  # [c, baseline, p]: xx and yy autos
    worst = 0.0
    for p1 in (0, 1):
        for p2 in (0, 1):
            norm = np.sqrt(np.maximum(auto[:, di, p1] * auto[:, dj, p2], 1e-300))
            worst = max(worst, float((dv[:, :, 2 * p1 + p2] / norm).max()))
    return {"max_abs_err": float(dv.max()),
            "max_abs_vis": float(mag.max()),
            "rms_rel_err": float(np.sqrt((dv ** 2).sum() / (mag ** 2).sum())),
            "max_norm_err": worst}
