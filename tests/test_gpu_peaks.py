"""alls_device_peaks on the GPU: the measured rates are positive and within
what the hardware can do (vendor peaks: 8 TB/s HBM, 78.6 / 157.3 TFLOP/s
dense FP64 / FP32 MFMA)."""
import pytest

import alls

pytestmark = pytest.mark.gpu


def test_device_peaks_are_plausible(ctx):
    pk = ctx.device_peaks()
    print("device peaks", pk, flush=True)
    assert 500.0 < pk["copy_GBs"] < 9000.0
    assert 5.0 < pk["mfma_f64_TFLOPs"] < 90.0
    assert 5.0 < pk["mfma_f32_TFLOPs"] < 180.0
    assert pk["atomic_f64_Gops"] > 1.0
