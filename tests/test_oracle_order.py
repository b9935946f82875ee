"""Order of accuracy of the oracle's AHL21 composition (ahl21.jl:5-95).

AHL21 is a 4th-order map: halving h divides the global error of a fixed
integration span by 2^4 = 16.  That order rests on the two correction
coefficients, h^3/36 * G for kicked pairs in phic! (ahl21.jl:401) and
alpha * h^3/96 * 2G for Kepler pairs in phisalpha! (ahl21.jl:564); any other
value leaves an O(h^3) error per step, i.e. a 2nd-order method (ratio 4).
The finite-difference, symplecticity, time-reversal and two-body checks hold
for any coefficient, so this is the test that pins them.  It also plants a
2 % error in each coefficient (an oracle test knob) and shows the ratio drop.
The GPU kernels equal the oracle bit for bit, so they inherit the pin.
"""
import numpy as np
import pytest

import oracle_lib as O

# star + three planets of a few 1e-3 stellar masses (strong mutual
# perturbations, so the correction terms matter), near the 2:1 chain
EL = np.array([[1.0, 0, 0, 0, 0, 0, 0],
               [3e-3, 10.0, 1.0, 0.05, 0.02, np.pi / 2 + 0.01, 0.0],
               [5e-3, 21.3, 4.0, -0.03, 0.04, np.pi / 2 - 0.02, 0.1],
               [2e-3, 47.0, 9.0, 0.02, -0.05, np.pi / 2 + 0.015, -0.05]])
SPAN = 160.0
H = (0.25, 0.125, 0.0625)
H_REF = 1.0 / 512


def _pair(kicked):
    if not kicked:
        return None
    p = np.zeros((4, 4), dtype=np.uint8)
    p[1, 2] = p[2, 1] = 1  # the two inner planets interact by kicks
    return p


def _x_after(h, pair):
    x, v, _ = O.init_nbody(0.0, EL)
    s = O.OState(x, v, EL[:, 0], t=0.0, pair=pair)
    O.integrate_steps(s, h, int(round(SPAN / h)), grad=False)
    return s.x.copy()


def _ratios(pair):
    ref = _x_after(H_REF, pair)
    err = np.array([np.max(np.abs(_x_after(h, pair) - ref)) for h in H])
    return err[:-1] / err[1:], err


@pytest.mark.parametrize("kicked", [False, True])
def test_fourth_order(kicked):
    r, err = _ratios(_pair(kicked))
    assert np.all(err > 1e-11), err  # truncation, well above rounding
    assert np.all((r > 14.0) & (r < 18.0)), (r, err)


@pytest.mark.parametrize("which", ["kicked", "kepler"])
def test_wrong_coefficient_is_second_order(which):
    # with the kicked pair present, a 2 % error in either coefficient drops
    # the error ratio to about 4
    try:
        O.set_correction_scale(kicked=1.02 if which == "kicked" else 1.0,
                               kepler=1.02 if which == "kepler" else 1.0)
        r, err = _ratios(_pair(True))
    finally:
        O.set_correction_scale(1.0, 1.0)
    assert r[-1] < 6.0, (r, err)
