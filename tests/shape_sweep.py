"""Seeded random shapes for the X-engine parity sweep.

One deterministic list per input type.  The GPU parity sweep
(test_gpu_parity.py::test_i8_shape_sweep / test_f16_shape_sweep) runs each
case against the oracle; test_dma_model.py::test_shape_sweep_in_bounds first
checks, on the CPU, every load and output store those cases make the kernels
issue.  A case is (ns, nf, nt, order, calls, tuning): nstation even in
2..402 (a multiple of 4 for register-tile order, which needs it, and for
most int8 cases, which the real-Gram kernels need), 1..5
channels, 1..3000 samples split over 1..3 calls (the later calls accumulate),
and one kernel variant of the engine's tuning.
"""
import numpy as np

# (kernel, staging, stage, mfma) of XGPUAmdTuning: the complex-plane kernel;
# the real-Gram kernel with register staging (int8 64 / 128, fp16 32 / 64
# samples per stage) and with global -> LDS DMA (int8 32 / 64, fp16 16 / 32); the
# int8 DMA kernel with 16x16x64 MFMA tiles; the K-interleaved pair-tile kernel
I8_TUNINGS = [(1, 0, 0, 0), (2, 1, 64, 0), (2, 1, 128, 0), (2, 2, 32, 0), (2, 2, 64, 0),
              (2, 2, 64, 2), (3, 0, 0, 0)]
F16_TUNINGS = [(1, 0, 0, 0), (2, 1, 32, 0), (2, 1, 64, 0), (2, 2, 16, 0), (2, 2, 32, 0)]
ORDER_NAMES = ("regtile", "tri", "realimag")


def _cases(kind, n, seed):
    rng = np.random.default_rng(seed)
    tunings = I8_TUNINGS if kind == "i8" else F16_TUNINGS
    out = []
    for k in range(n):
        order = ORDER_NAMES[int(rng.integers(3))]
        ns = 2 * int(rng.integers(1, 201))
        # int8: mostly nstation % 4 == 0, which the real-Gram kernels take
        if (order == "regtile" or (kind == "i8" and rng.random() < 0.8)) and ns % 4:
            ns += 2
        nf = int(rng.integers(1, 6))
        # half the blocks short (time tails inside the first stages), half long
        nt = int(rng.integers(1, 300)) if rng.random() < 0.5 else int(rng.integers(300, 3001))
        calls = min(int(rng.integers(1, 4)), nt)
        kernel, staging, stage, mfma = tunings[k % len(tunings)]
        out.append((ns, nf, nt, order, calls,
                    dict(kernel=kernel, staging=staging, stage=stage, mfma=mfma)))
    return out


I8_SWEEP = _cases("i8", 35, 20261)
F16_SWEEP = _cases("f16", 20, 20262)


def call_lengths(nt, calls):
    """Samples per call: nt // calls each, the remainder in the last."""
    per = nt // calls
    return [per] * (calls - 1) + [nt - per * (calls - 1)]


def case_id(case):
    ns, nf, nt, order, calls, t = case
    return (f"ns{ns}-nf{nf}-nt{nt}-{order}-c{calls}-k{t['kernel']}s{t['staging']}"
            f"g{t['stage']}m{t['mfma']}")
