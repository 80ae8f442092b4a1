"""Pin the CPU oracle (oracle/xgpu_oracle.c) before trusting it.

Checks: golden vectors of the compiled reference (tests/golden), numpy brute
force from the definition (tests/brute.py), closed forms and invariants.  A
planted mistake in the oracle (dropped term, sign, transposed operand, wrong
index) fails at least one of these.
"""
import hashlib

import numpy as np
import pytest

import oracle
from brute import fill_uniform, pack_tri, regtile_index, xengine_exact, xengine_f64
from conftest import golden


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- generators
def test_fill_uniform_matches_numpy_restatement():
    for seed, lo, hi in ((0, -128, 127), (7, -127, 127), (123, -5, 9), (2**40 + 3, 0, 0)):
        a = oracle.fill_uniform(4099, seed, lo, hi)
        b = fill_uniform(4099, seed, lo, hi)
        assert np.array_equal(a, b)
        assert a.min() >= lo and a.max() <= hi


def test_fill_uniform_covers_range_uniformly():
    a = oracle.fill_uniform(256 * 4000, 5, -128, 127).astype(np.int64)
    counts = np.bincount(a + 128, minlength=256)
    assert counts.min() > 3400 and counts.max() < 4600  # includes -128
    assert counts[0] > 0


def test_random_legacy_matches_reference(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    m = meta["ns16_nf2_nt64"]
    ncomplex = m["nf"] * m["nt"] * m["ns"] * 2
    x = oracle.random_legacy(ncomplex, m["srand_seed"])
    assert np.array_equal(x, g["x_legacy"])
    vals = np.unique(x)
    assert set(vals.tolist()) <= set(range(-112, 113, 16))


def test_random_legacy_hash_ns32(meta):
    m = meta["ns32_nf3_nt256"]
    x = oracle.random_legacy(m["nf"] * m["nt"] * m["ns"] * 2, m["srand_seed"])
    assert sha(x) == m["x_sha256"]


# ---------------------------------------------------------------- float X-engine
def test_xengine_f32_matches_reference_golden(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    m = meta["ns16_nf2_nt64"]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    assert np.array_equal(oracle.xengine_f32(g["x_legacy"], ns, nf, nt), g["omp_legacy"])
    assert np.array_equal(oracle.xengine_f32(g["x_full"], ns, nf, nt), g["omp_full"])


def test_xengine_f32_matches_reference_ns32(meta):
    m = meta["ns32_nf3_nt256"]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.random_legacy(nf * nt * ns * 2, m["srand_seed"])
    assert np.array_equal(oracle.xengine_f32(x, ns, nf, nt), golden("ref_ns32_nf3_nt256.npz")["omp"])


@pytest.mark.parametrize("tag", ["ns256_nf10_nt1024", "ns512_nf1_nt8192"])
def test_xengine_f32_full_size_hash(meta, tag):
    """Bit-identical to the reference at the cfg1 shape and at one 512-station
    channel of 8192 samples (float rounding past 2^24 included)."""
    m = meta[tag]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    assert sha(x) == m["x_sha256"]
    out = oracle.xengine_f32(x, ns, nf, nt)
    assert sha(out) == m["omp_sha256"]


def test_dp4a_float_accumulation_matches_reference(meta):
    """The reference -DDP4A build rounds int sums through float past 2^24."""
    m = meta["ns16_nf2_nt8192_dp4a"]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    assert sha(x) == m["x_sha256"]
    ref = golden("ref_ns16_nf2_nt8192_dp4a.npz")["omp"]
    assert np.array_equal(oracle.xengine_dp4a_f32acc(x, ns, nf, nt), ref)
    exact = oracle.xengine_i64(x, ns, nf, nt)
    diff = exact != ref.astype(np.int64)
    assert diff.any(), "NT=8192 full-range autocorrelations exceed 2^24"
    # only sums beyond 2^24 may differ
    assert np.all(np.abs(exact[diff]) > 2**24)


def test_dp4a_reference_exact_below_2_24(meta):
    m = meta["ns48_nf5_nt512_dp4a"]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    assert sha(x) == m["x_sha256"]
    ref = golden("ref_ns48_nf5_nt512_dp4a.npz")["omp"].astype(np.int64)
    assert np.array_equal(oracle.xengine_i64(x, ns, nf, nt), ref)
    assert np.array_equal(oracle.xengine_dp4a_f32acc(x, ns, nf, nt).astype(np.int64), ref)


# ---------------------------------------------------------------- exact X-engine
@pytest.mark.parametrize("ns,nf,nt,seed", [(2, 1, 5, 1), (6, 3, 17, 2), (16, 2, 64, 3), (10, 1, 200, 4)])
def test_xengine_i64_brute_force(ns, nf, nt, seed):
    x = oracle.fill_uniform(nt * nf * ns * 4, seed, -128, 127)
    assert np.array_equal(oracle.xengine_i64(x, ns, nf, nt), xengine_exact(x, ns, nf, nt))


def test_xengine_i64_long_blocks_brute_force():
    """Blocks longer than the oracle's int32 partial-sum chunk (16384
    samples), at -128 everywhere (the largest per-sample terms) and on
    uniform data."""
    ns, nf, nt = 2, 1, 40000
    x = np.full(nt * nf * ns * 4, -128, dtype=np.int8)
    v = oracle.xengine_i64(x, ns, nf, nt).reshape(-1, 2)
    assert np.all(v[:, 0] == nt * 2 * 128 * 128) and np.all(v[:, 1] == 0)
    x = oracle.fill_uniform(nt * nf * ns * 4, 31, -128, 127)
    assert np.array_equal(oracle.xengine_i64(x, ns, nf, nt), xengine_exact(x, ns, nf, nt))


def test_xengine_i64_channels_equals_full():
    ns, nf, nt = 20, 5, 300
    x = oracle.fill_uniform(nt * nf * ns * 4, 32, -128, 127)
    full = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 4, 2)
    for chans in ([0], [4, 1], [2, 2, 3]):
        assert np.array_equal(oracle.xengine_i64_channels(x, ns, nf, nt, chans), full[chans])


def test_xengine_i64_closed_forms():
    ns, nf, nt = 8, 2, 33
    # constant input c = a + bi on every station/pol: V = NT |c|^2 (imag 0)
    for a, b in ((3, -4), (-128, -128), (127, -128)):
        x = np.tile(np.array([a, b], dtype=np.int8), nt * nf * ns * 2)
        v = oracle.xengine_i64(x, ns, nf, nt).reshape(-1, 2)
        assert np.all(v[:, 0] == nt * (a * a + b * b))
        assert np.all(v[:, 1] == 0)
    # -128 - 128i against 1 + 0i: re = -128, im = -(-128)*... exact sign check
    x = np.zeros((nt, nf, ns, 2, 2), dtype=np.int8)
    x[:, :, 3, 0] = (-128, -128)
    x[:, :, 1, 1] = (1, 0)
    v = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 2, 2, 2)
    k = 3 * 4 // 2 + 1  # (i=3, j=1)
    # V = sum x3 * conj(x1) = nt * (-128 - 128i) * 1
    assert v[0, k, 0, 1, 0] == -128 * nt and v[0, k, 0, 1, 1] == -128 * nt
    # a single nonzero station: only its autocorrelations are nonzero
    x = np.zeros((nt, nf, ns, 2, 2), dtype=np.int8)
    x[:, :, 5] = oracle.fill_uniform(nt * nf * 4, 9, -128, 127).reshape(nt, nf, 2, 2)
    v = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 8)
    nz = np.nonzero(np.any(v != 0, axis=(0, 2)))[0]
    assert nz.tolist() == [5 * 6 // 2 + 5]


def test_xengine_i64_invariants():
    ns, nf, nt = 12, 2, 77
    x = oracle.fill_uniform(nt * nf * ns * 4, 21, -128, 127)
    v = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 2, 2, 2)
    for i in range(ns):
        k = i * (i + 1) // 2 + i
        for p in range(2):
            assert np.all(v[:, k, p, p, 1] == 0)       # autos real
            assert np.all(v[:, k, p, p, 0] >= 0)       # and non-negative
        # XY and YX of an auto baseline are conjugates
        assert np.all(v[:, k, 0, 1, 0] == v[:, k, 1, 0, 0])
        assert np.all(v[:, k, 0, 1, 1] == -v[:, k, 1, 0, 1])
    # linearity in time: two halves add up
    a = oracle.xengine_i64(x[: 40 * nf * ns * 4], ns, nf, 40)
    b = oracle.xengine_i64(x[40 * nf * ns * 4:], ns, nf, nt - 40)
    assert np.array_equal(a + b, v.reshape(-1))


def test_xengine_i64_pairs_agree():
    ns, nf, nt = 20, 3, 50
    x = oracle.fill_uniform(nt * nf * ns * 4, 4, -128, 127)
    full = oracle.xengine_i64(x, ns, nf, nt).reshape(nf, -1, 4, 2)
    rng = np.random.default_rng(0)
    fij = []
    for _ in range(40):
        f = rng.integers(nf)
        i = rng.integers(ns)
        j = rng.integers(i + 1)
        fij.append((f, i, j))
    got = oracle.xengine_i64_pairs(x, ns, nf, nt, fij)
    for n, (f, i, j) in enumerate(fij):
        assert np.array_equal(got[n], full[f, i * (i + 1) // 2 + j])


def test_xengine_f32_exact_below_2_24():
    ns, nf, nt = 14, 2, 300
    x = oracle.fill_uniform(nt * nf * ns * 4, 8, -127, 127)
    assert np.array_equal(oracle.xengine_f32(x, ns, nf, nt).astype(np.int64),
                          oracle.xengine_i64(x, ns, nf, nt))


# ---------------------------------------------------------------- fp16 oracle
def test_xengine_f64_f16_brute_force():
    ns, nf, nt = 6, 2, 40
    rng = np.random.default_rng(1)
    xc = (rng.normal(0, 16, (nt, nf, ns, 2)) + 1j * rng.normal(0, 16, (nt, nf, ns, 2)))
    x16 = np.stack([xc.real, xc.imag], axis=-1).astype(np.float16)
    exact = x16.astype(np.float64)
    ref = xengine_f64(exact[..., 0] + 1j * exact[..., 1], ns, nf, nt)
    got = oracle.xengine_f64_f16(x16, ns, nf, nt)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_half_decoding_special_values():
    vals = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 6.1e-5, 5.96e-8, -1e-7], dtype=np.float16)
    ns, nf, nt = 2, 1, len(vals)
    x16 = np.zeros((nt, nf, ns, 2, 2), dtype=np.float16)
    x16[:, 0, 0, 0, 0] = vals
    x16[:, 0, 0, 0, 1] = 1.0
    got = oracle.xengine_f64_f16(x16, ns, nf, nt).reshape(-1, 2)
    v = vals.astype(np.float64)
    assert got[0, 0] == pytest.approx(float(np.sum(v * v + 1.0)), rel=1e-15)


# ---------------------------------------------------------------- layouts
def test_reorder_regtile_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    ns, nf = 16, 2
    tri = oracle.reorder(3000, g["regtile"], ns, nf)
    n = 2 * oracle.tri_len(ns, nf)
    assert np.array_equal(tri, g["reordered"][:n])


def test_reorder_realimag_matches_reference():
    g = golden("ref_ns16_nf2_nt64_ri.npz")
    tri = oracle.reorder(2000, g["realimag"], 16, 2)
    assert np.array_equal(tri, g["reordered"])


def test_regtile_roundtrip_and_formula():
    ns, nf = 12, 3
    tri = np.arange(2 * oracle.tri_len(ns, nf), dtype=np.int64) + 7
    reg = oracle.regtile_from_tri(tri, ns, nf)
    assert np.array_equal(oracle.reorder(3000, reg, ns, nf), tri)
    plane = oracle.regtile_len(ns, nf)
    nb = ns * (ns + 1) // 2
    for f, i, j, p1, p2 in ((0, 0, 0, 0, 0), (1, 5, 2, 1, 0), (2, 11, 11, 1, 1), (2, 7, 6, 0, 1)):
        k = f * nb + i * (i + 1) // 2 + j
        idx = regtile_index(f, i, j, p1, p2, ns)
        assert reg[idx] == tri[2 * ((k * 2 + p1) * 2 + p2)]
        assert reg[idx + plane] == tri[2 * ((k * 2 + p1) * 2 + p2) + 1]
    # quadrant "12" of diagonal tiles is never written
    written = np.zeros(2 * plane, bool)
    for f in range(nf):
        for i in range(ns):
            for j in range(i + 1):
                for p in range(4):
                    written[regtile_index(f, i, j, p // 2, p % 2, ns)] = True
    assert np.all(reg[:plane][~written[:plane]] == 0)
    assert (~written[:plane]).sum() == nf * (ns // 2) * 4


def test_swizzle_matches_reference(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    assert np.array_equal(oracle.swizzle(g["x_legacy"], 16, 2, 64), g["swizzled"])
    assert meta["ns16_nf2_nt64"]["swizzle_text"] == "Swizzling input\n"


@pytest.mark.parametrize("tag", ["ns16_nf2_nt64_cb32", "ns32_nf3_nt256_cb32_dp4a"])
def test_unblock32_matches_reference_block32_build(meta, tag):
    """The reference CPU engine built -DCOMPLEX_BLOCK_SIZE=32 (float and DP4A)
    equals the exact engine on the unblocked input (sums stay below 2^24, so
    the float build is exact too); swapping the real and imaginary halves of
    each block, or leaving the input blocked, does not."""
    m = meta[tag]
    ns, nf, nt = m["ns"], m["nf"], m["nt"]
    x = oracle.fill_uniform(nt * nf * ns * 4, m["x_seed"], *m["x_range"])
    assert sha(x) == m["x_sha256"]
    ref = golden(f"ref_{tag}.npz")["omp"].astype(np.int64)
    assert np.array_equal(oracle.xengine_i64(oracle.unblock32(x), ns, nf, nt), ref)
    swapped = x.reshape(-1, 2, 32)[:, ::-1, :].reshape(-1)
    assert not np.array_equal(oracle.xengine_i64(oracle.unblock32(swapped), ns, nf, nt), ref)
    assert not np.array_equal(oracle.xengine_i64(x, ns, nf, nt), ref)


def test_unblock32_layout_and_inverse():
    """Byte q of a 64-byte block is the real part of plain entry q, byte
    32 + q its imaginary part; block32 inverts unblock32."""
    x = np.arange(128, dtype=np.int64).astype(np.int8)
    u = oracle.unblock32(x)
    for b in (0, 64):
        for q in range(32):
            assert u[b + 2 * q] == x[b + q] and u[b + 2 * q + 1] == x[b + 32 + q]
    y = oracle.fill_uniform(64 * 37, 3, -128, 127)
    assert np.array_equal(oracle.block32(oracle.unblock32(y)), y)
    assert np.array_equal(oracle.unblock32(oracle.block32(y)), y)


def test_extract_matches_reference():
    g = golden("ref_ns16_nf2_nt64.npz")
    assert np.array_equal(oracle.extract(g["omp_legacy"], 16, 2), g["full"])


def test_check_counts_match_reference_report(meta):
    g = golden("ref_ns16_nf2_nt64.npz")
    n, mx = oracle.check(g["check_gpu"], g["omp_legacy"], 16, 2)
    text = meta["ns16_nf2_nt64"]["check_text"]["0"]
    assert f"failed with {n} deviations (max error {mx:g})" in text
    n0, mx0 = oracle.check(g["omp_legacy"], g["omp_legacy"], 16, 2)
    assert n0 == 0 and mx0 == 0


def test_pack_tri_helper_consistent_with_oracle_layout():
    ns, nf, nt = 4, 1, 3
    x = oracle.fill_uniform(nt * nf * ns * 4, 1, -128, 127)
    a = np.asarray(x).reshape(nt, nf, ns * 2, 2).astype(np.int64)
    re = np.einsum("tfm,tfn->fmn", a[..., 0], a[..., 0]) + np.einsum("tfm,tfn->fmn", a[..., 1], a[..., 1])
    im = np.einsum("tfm,tfn->fmn", a[..., 1], a[..., 0]) - np.einsum("tfm,tfn->fmn", a[..., 0], a[..., 1])
    assert np.array_equal(pack_tri(re, im, ns, nf), oracle.xengine_i64(x, ns, nf, nt))
