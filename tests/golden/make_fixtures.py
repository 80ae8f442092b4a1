"""Write the golden vectors under tests/golden from the reference's own CPU code.

The reference's src/omp_xengine.c and src/cpu_util.c are compiled as they lie
by oracle/Makefile into oracle/_ref/libxgpuref_<sizing>.so (one library per
build-time sizing).  This script calls them and stores inputs (or the seed of
the generator that makes them) and outputs as .npz / .json fixtures.  Run it
in a container that has /root/reference:

    make -C oracle && python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

libc = C.CDLL("libc.so.6")


def ref(tag):
    L = oracle.ref_lib(tag)
    if L is None:
        raise SystemExit(f"oracle/_ref/libxgpuref_{tag}.so missing: run make -C oracle")
    return L


def capture_stdout(fn):
    """Run fn() with the process's fd 1 redirected; return the text printed."""
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


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tri_len(ns, nf):
    return nf * ns * (ns + 1) // 2 * 4


def regtile_len(ns, nf):
    return nf * (ns // 2 + 1) * (ns // 4) * 16


def omp(L, x, ns, nf, dtype=np.float32):
    out = np.zeros(2 * tri_len(ns, nf), dtype=dtype)
    L.xgpuOmpXengine(out.ctypes.data, np.ascontiguousarray(x).ctypes.data)
    return out


def main():
    meta = {}
    # --- 1. NS=16, NF=2, NT=64, default (register tile, float) ------------
    ns, nf, nt = 16, 2, 64
    L = ref("ns16_nf2_nt64")
    veclen = nf * nt * ns * 2
    libc.srand(1)
    x_legacy = np.zeros(2 * veclen, dtype=np.int8)
    L.xgpuRandomComplex(x_legacy.ctypes.data, veclen)
    x_full = oracle.fill_uniform(2 * veclen, 7, -128, 127)
    omp_legacy = omp(L, x_legacy, ns, nf)
    omp_full = omp(L, x_full, ns, nf)
    # register-tile buffer with distinct values -> reorder
    mat = regtile_len(ns, nf)
    reg = (np.arange(2 * mat, dtype=np.float32) * 3 + 1).astype(np.float32)
    reordered = reg.copy()
    L.xgpuReorderMatrix(reordered.ctypes.data)
    full = np.zeros(nf * ns * ns * 4 * 2, dtype=np.float32)
    L.xgpuExtractMatrix(full.ctypes.data, omp_legacy.ctypes.data)
    swz = np.zeros_like(x_legacy)
    swz_text = capture_stdout(lambda: L.xgpuSwizzleInput(swz.ctypes.data, x_legacy.ctypes.data))
    # xgpuCheckResult on planted errors (it rounds its first argument in place)
    gpu = omp_legacy.copy()
    gpu[10] += 3.0          # relative error on one visibility
    gpu[2 * 17 + 1] += 0.4  # rounds away
    gpu[2 * 300] = -gpu[2 * 300] + 1.0
    gpu_in = gpu.copy()
    check_text = {}
    for verbose in (0, 1):
        g = gpu_in.copy()
        check_text[str(verbose)] = capture_stdout(
            lambda: L.xgpuCheckResult(g.ctypes.data, omp_legacy.ctypes.data, verbose, None))
    g = omp_legacy.copy()
    check_text["ok"] = capture_stdout(
        lambda: L.xgpuCheckResult(g.ctypes.data, omp_legacy.ctypes.data, 0, None))
    np.savez_compressed(
        os.path.join(HERE, "ref_ns16_nf2_nt64.npz"), x_legacy=x_legacy, x_full=x_full,
        omp_legacy=omp_legacy, omp_full=omp_full, regtile=reg, reordered=reordered, full=full,
        swizzled=swz, check_gpu=gpu_in)
    meta["ns16_nf2_nt64"] = {"ns": ns, "nf": nf, "nt": nt, "x_full_seed": 7,
                             "x_full_range": [-128, 127], "srand_seed": 1,
                             "swizzle_text": swz_text, "check_text": check_text}

    # --- 2. REAL_IMAG reorder ------------------------------------------------
    L = ref("ns16_nf2_nt64_ri")
    mat = tri_len(ns, nf)
    ri = (np.arange(2 * mat, dtype=np.float32) * 5 + 2).astype(np.float32)
    ri_out = ri.copy()
    L.xgpuReorderMatrix(ri_out.ctypes.data)
    np.savez_compressed(os.path.join(HERE, "ref_ns16_nf2_nt64_ri.npz"), realimag=ri,
                        reordered=ri_out)

    # --- 3. DP4A build: int Complex accumulated through float --------------
    ns, nf, nt = 16, 2, 8192
    L = ref("ns16_nf2_nt8192_dp4a")
    x = oracle.fill_uniform(nt * nf * ns * 4, 3, -128, 127)
    out = omp(L, x, ns, nf, np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_ns16_nf2_nt8192_dp4a.npz"), omp=out)
    meta["ns16_nf2_nt8192_dp4a"] = {"ns": ns, "nf": nf, "nt": nt, "x_seed": 3,
                                    "x_range": [-128, 127], "x_sha256": sha(x)}

    # --- 4. NS=32, NF=3, NT=256, legacy data ----------------------------------
    ns, nf, nt = 32, 3, 256
    L = ref("ns32_nf3_nt256")
    veclen = nf * nt * ns * 2
    libc.srand(5)
    x = np.zeros(2 * veclen, dtype=np.int8)
    L.xgpuRandomComplex(x.ctypes.data, veclen)
    out = omp(L, x, ns, nf)
    np.savez_compressed(os.path.join(HERE, "ref_ns32_nf3_nt256.npz"), omp=out)
    meta["ns32_nf3_nt256"] = {"ns": ns, "nf": nf, "nt": nt, "srand_seed": 5, "x_sha256": sha(x)}

    # --- 5. NS=48, NF=5, NT=512, DP4A, full range (sums below 2^24) -----------
    ns, nf, nt = 48, 5, 512
    L = ref("ns48_nf5_nt512_dp4a")
    x = oracle.fill_uniform(nt * nf * ns * 4, 11, -128, 127)
    out = omp(L, x, ns, nf, np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_ns48_nf5_nt512_dp4a.npz"), omp=out)
    meta["ns48_nf5_nt512_dp4a"] = {"ns": ns, "nf": nf, "nt": nt, "x_seed": 11,
                                   "x_range": [-128, 127], "x_sha256": sha(x)}

    # --- 6. full-size hashes: cfg1 shape and one 512-station channel ----------
    for tag, (ns, nf, nt), seed, rng in (
            ("ns256_nf10_nt1024", (256, 10, 1024), 1, (-127, 127)),
            ("ns512_nf1_nt8192", (512, 1, 8192), 2, (-128, 127))):
        L = ref(tag)
        x = oracle.fill_uniform(nt * nf * ns * 4, seed, *rng)
        out = omp(L, x, ns, nf)
        meta[tag] = {"ns": ns, "nf": nf, "nt": nt, "x_seed": seed, "x_range": list(rng),
                     "x_sha256": sha(x), "omp_sha256": sha(out),
                     "omp_sum_re": float(out[0::2].astype(np.float64).sum()),
                     "omp_max": float(np.abs(out).max())}

    # --- 7. COMPLEX_BLOCK_SIZE=32 input layout (float and DP4A builds) -------
    for tag, (ns, nf, nt), seed, dtype in (
            ("ns16_nf2_nt64_cb32", (16, 2, 64), 13, np.float32),
            ("ns32_nf3_nt256_cb32_dp4a", (32, 3, 256), 17, np.int32)):
        L = ref(tag)
        x = oracle.fill_uniform(nt * nf * ns * 4, seed, -128, 127)  # block-32 bytes
        out = omp(L, x, ns, nf, dtype)
        np.savez_compressed(os.path.join(HERE, f"ref_{tag}.npz"), omp=out)
        meta[tag] = {"ns": ns, "nf": nf, "nt": nt, "x_seed": seed, "x_range": [-128, 127],
                     "x_sha256": sha(x), "layout": "COMPLEX_BLOCK_SIZE=32"}

    with open(os.path.join(HERE, "ref_meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
