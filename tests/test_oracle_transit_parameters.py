"""TransitParameters derivatives of the oracle against binary128 finite
differences: test_transit_parameters.jl:46-87 restated (3 bodies of the
TRAPPIST-1 fixture, masses x10, h = 0.04, 10 days; each element varied by
+-1e-10 with the mass last, both runs without gradients).

The reference's set-up is edge-on with Omega = 0, where b^2 and its
derivatives nearly vanish; here the two planets are inclined by ~0.1 rad,
eccentric (e ~ 0.1) and have rotated nodes, so all three rows -- the times,
v_sky and b^2 -- carry O(1) derivatives, and the v_sky row's dv_sky/dt * dt/dq
term (timing.jl:179-184) is about 1 % of that row.

Rows 0 and 1 agree with the finite differences to ~1e-13 of each row's
largest entry.  Row 2 (b^2) agrees to ~4e-8: timing.jl:177 takes
d(b^2)/dt = 2 g_sky = 0 at the transit, which holds for the exact flow, while
the step map's own dx/dh differs from v by the method's truncation error
(the gap falls from 1.9e-6 at h = 0.08 to 1.9e-8 at h = 0.02).  The
reference's own criterion (asinh, maxabs over every entry, rtol = sqrt(eps))
is met by many orders of magnitude.
"""
import numpy as np
import pytest

import oracle_lib as O
from refcases import elements


def _case():
    el = elements()[:3].copy()
    el[1:, 2] -= 7300.0
    el[1, 0] *= 10.0
    el[2, 0] *= 10.0
    el[1, 3:7] = [0.08, 0.06, np.pi / 2 - 0.12, 0.25]
    el[2, 3:7] = [-0.05, 0.09, np.pi / 2 + 0.09, -0.3]
    return el, 7257.93115525 - 7300.0 - 0.5, 0.04, 10.0


@pytest.fixture(scope="module")
def run():
    el, t0, h, tmax = _case()
    ntt = O.ntt_rule(tmax, el)
    x, v, ji = O.init_nbody(t0, el)
    s = O.OState(x, v, el[:, 0], t=t0, jac_init=ji)
    ttbv, dq, de, count = O.integrate_transits(s, h, tmax, ntt, ti=0, grad=True, ntbv=3)
    fd = O.fd_dtbvdelements(t0, el, h, tmax, 0, ntt, ntbv=3)
    return el, t0, h, x, v, ji, ttbv, de, count, fd


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_every_row_against_finite_differences(run):
    el, t0, h, x, v, ji, ttbv, de, count, fd = run
    assert count[1] >= 5 and count[2] >= 3
    assert _rel(de[0], fd[0]) < 1e-11  # transit times
    assert _rel(de[1], fd[1]) < 1e-11  # v_sky
    assert _rel(de[2], fd[2]) < 1e-6   # b^2 (see the module docstring)
    # test_transit_parameters.jl:86-87
    assert np.max(np.abs(np.arcsinh(de) - np.arcsinh(fd))) / np.max(np.abs(np.arcsinh(fd))) < 1.5e-8
    # the b^2 rows are not degenerate in this set-up
    assert np.max(np.abs(fd[2])) > 1e-4 and np.min(ttbv[2, 1:, :3]) > 1e-6


def test_vsky_time_term_is_resolved(run):
    """A v_sky row without its dv_sky/dt * dt/dq term would miss the finite
    differences by far more than the tolerance above: dv_sky/dt at each
    transit, from v_sky at t +- 1e-4 d along the trajectory, times the time
    row, is ~1 % of the v_sky row."""
    el, t0, h, x, v, ji, ttbv, de, count, fd = run
    big = 0.0
    for i in (1, 2):
        k = int(count[i])
        times = ttbv[0, i, :k]
        sp = O.OState(x, v, el[:, 0], t=t0, jac_init=ji)
        vp, _, _, _ = O.transit_snapshot(sp, h, times + 1e-4, grad=False)
        sm = O.OState(x, v, el[:, 0], t=t0, jac_init=ji)
        vm, _, _, _ = O.transit_snapshot(sm, h, times - 1e-4, grad=False)
        dvdt = (vp[i] - vm[i]) / 2e-4
        big = max(big, np.max(np.abs(dvdt[:, None, None] * de[0, i, :k])))
    assert big > 1e6 * 1e-11 * np.max(np.abs(fd[1])), big
