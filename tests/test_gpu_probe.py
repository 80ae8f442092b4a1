"""The MFMA-only probes (include/xgpu_amd.h: xgpuAmdMfmaPeak,
xgpuAmdMfmaPeakData): positive rates below the dense int8 peak.  The
operand-driven probe is the ceiling bench.py compares the X-engine with on
uniform full-range input."""
import numpy as np
import pytest

import xgpu_amd as X
from conftest import flush

pytestmark = pytest.mark.gpu

DENSE_PEAK_CMAC_S = 5.0e15 / 8   # 5.0 POP/s dense int8, 8 int ops per complex MAC


def test_mfma_probes_report_rates_below_the_dense_peak():
    flush("[probe] mfma peak, constant operands")
    const = X.mfma_peak(reps=2)
    assert 0.2 * DENSE_PEAK_CMAC_S < const < 1.05 * DENSE_PEAK_CMAC_S
    ops = X.DeviceBuffer(X.PEAK_OPERAND_BYTES)
    ops.fill_uniform_int8(5, -128, 127)
    flush("[probe] mfma peak, uniform operands")
    uni = X.mfma_peak_data(ops, reps=2)
    ops.upload(np.full(X.PEAK_OPERAND_BYTES, 3, np.int8))
    flush("[probe] mfma peak, constant operand buffer")
    three = X.mfma_peak_data(ops, reps=2)
    ops.free()
    assert 0.1 * DENSE_PEAK_CMAC_S < uni < 1.05 * DENSE_PEAK_CMAC_S
    assert 0.1 * DENSE_PEAK_CMAC_S < three < 1.05 * DENSE_PEAK_CMAC_S


def test_fp16_mfma_ceiling_probe():
    """fp16 operands (Gaussian, sigma 16) on the fp16 MFMA: a positive rate
    below the dense FP16 peak (2.5 PFLOP/s = 312.5 T complex MACs/s)."""
    g = (np.random.default_rng(3).standard_normal(X.PEAK_OPERAND_BYTES // 2) * 16).astype(np.float16)
    ops = X.DeviceBuffer.from_array(g)
    flush("[probe] fp16 mfma peak, Gaussian operands")
    rate = X.mfma_peak_data(ops, reps=2, compute=X.XGPU_FLOAT16)
    ops.free()
    assert 0.1 * 312.5e12 < rate < 1.05 * 312.5e12
