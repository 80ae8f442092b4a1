"""numpy brute-force X-engine and layout maps, written from the definitions.

Independent of oracle/xgpu_oracle.c: used to pin the C oracle on small inputs.
V[f,i,j,p1,p2] = sum_t x[t,f,i,p1] * conj(x[t,f,j,p2]) for stations i >= j,
triangular index k = f*NB + i(i+1)/2 + j, NB = NS(NS+1)/2.
"""
import numpy as np


def as_complex(x_int8, ns, nf, nt):
    a = np.asarray(x_int8).reshape(nt, nf, ns, 2, 2).astype(np.int64)
    return a[..., 0] + 1j * a[..., 1]  # complex128, exact for these magnitudes


def xengine_exact(x_int8, ns, nf, nt):
    """int64 (re, im) interleaved, triangular order."""
    a = np.asarray(x_int8).reshape(nt, nf, ns * 2, 2).astype(np.int64)
    ar, ai = a[..., 0], a[..., 1]
    # full (2ns x 2ns) matrices per channel: [f, m, n]
    re = np.einsum("tfm,tfn->fmn", ar, ar) + np.einsum("tfm,tfn->fmn", ai, ai)
    im = np.einsum("tfm,tfn->fmn", ai, ar) - np.einsum("tfm,tfn->fmn", ar, ai)
    return pack_tri(re, im, ns, nf)


def pack_tri(re, im, ns, nf):
    """[f, m, n] full matrices (m = 2i+p1 row, n = 2j+p2 column) -> triangular."""
    nb = ns * (ns + 1) // 2
    out = np.zeros((nf, nb, 2, 2, 2), dtype=re.dtype)
    for i in range(ns):
        for j in range(i + 1):
            k = i * (i + 1) // 2 + j
            for p1 in range(2):
                for p2 in range(2):
                    out[:, k, p1, p2, 0] = re[:, 2 * i + p1, 2 * j + p2]
                    out[:, k, p1, p2, 1] = im[:, 2 * i + p1, 2 * j + p2]
    return out.reshape(-1)


def xengine_f64(xc, ns, nf, nt):
    """complex input [t,f,s,p] (complex128) -> interleaved float64 triangular."""
    a = xc.reshape(nt, nf, ns * 2)
    v = np.einsum("tfm,tfn->fmn", a, np.conj(a))
    return pack_tri(v.real, v.imag, ns, nf)


def regtile_index(f, i, j, p1, p2, ns):
    """Real-plane index of visibility (f,i>=j,p1,p2) in REGISTER_TILE order."""
    P = ns // 2
    RT = P * (P + 1) // 2
    R, C = i // 2, j // 2
    q = 2 * (j % 2) + (i % 2)
    l_ = (f * 4 + q) * RT + R * (R + 1) // 2 + C
    return (l_ * 2 + p1) * 2 + p2


def mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def fill_uniform(n, seed, lo, hi):
    """Counter-based uniform bytes, vectorised restatement of the generator."""
    key = np.uint64(mix64(seed))
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = key + i + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    span = np.uint64(hi - lo + 1)
    v = ((x >> np.uint64(32)) * span) >> np.uint64(32)
    return (v.astype(np.int64) + lo).astype(np.int8)
