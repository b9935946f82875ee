"""The row-update probe (nbg_probe_row_update): its VALU and FP64-MFMA forms
compute the same product (the matrix-core lane layout of
v_mfma_f64_16x16x4f64 is as the kernel assumes) to rounding, and a
reference computed here in numpy agrees with both."""
import numpy as np
import pytest

from nbodygradient_amd import probe

pytestmark = pytest.mark.gpu


def _reference(reps, w=0):
    # the probe's synthetic data (nbg_probe.hip: coef, jinit) and update,
    # with numpy's float64 arithmetic and an ordinary sum
    r = np.arange(12)[:, None]
    c = np.arange(63)[None, :]
    J = 1.0 + 0.01 * ((13 * r + 7 * c + w) % 17)
    q = np.arange(12)[:, None]
    k = np.arange(14)[None, :]
    Cs = [2e-5 * ((7 * q + 3 * k + s) % 11 - 5) for s in (0, 1)]
    e6 = (np.arange(63) == 6).astype(float)
    e13 = (np.arange(63) == 13).astype(float)
    for it in range(reps):
        B = np.vstack([J[:6], e6, J[6:], e13])
        J = J + Cs[it & 1] @ B
    return J


def test_valu_and_mfma_forms_agree():
    print("gpu: row-update probe", flush=True)
    reps = 300
    _, v = probe.run(0, reps, 4)
    _, m = probe.run(1, reps, 4)
    ref = _reference(reps)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(v - ref)) / scale < 1e-12
    assert np.max(np.abs(m - ref)) / scale < 1e-12
    assert np.max(np.abs(v - m)) / scale < 1e-12
