"""Pin the oracle's transit search and gradients (timing.jl, Transits.jl).

Restates test/test_transit_timing.jl (dtdelements against binary128 finite
differences of all elements, every transited body ti; bitwise equal times
with and without gradients), test/test_findtransit.jl (recovery of the input
transit times) and adds two-body closed forms (transits spaced by exactly P,
the first one at the input t0 when I = pi/2).
"""
import numpy as np
import pytest

import oracle_lib as O
from refcases import elements, rel_maxabs, transit_case


def _run(el, t0, h, tmax, ti=0, grad=True, ntbv=1):
    x, v, jac_init = O.init_nbody(t0, el)
    s = O.OState(x, v, el[:, 0], t=t0, jac_init=jac_init)
    ntt = O.ntt_rule(tmax, el)
    tt, dtdq0, dtdel, count = O.integrate_transits(s, h, tmax, ntt, ti=ti, grad=grad, ntbv=ntbv)
    return s, tt, dtdq0, dtdel, count, ntt


@pytest.mark.parametrize("ti", [0, 1, 2])
def test_dtdelements_vs_finite_differences(ti):
    el, t0 = transit_case()
    h, itime = 0.04, 10.0
    s, tt, dtdq0, dtdel, count, ntt = _run(el, t0, h, itime, ti=ti)
    assert count.sum() > 0
    num = O.fd_dtdelements(t0, el, h, itime, ti, ntt, dq0=1e-10)
    # test_transit_timing.jl:87-88: asinh + max-abs at rtol = sqrt(eps)
    assert rel_maxabs(np.arcsinh(dtdel), np.arcsinh(num)) < 1e-9
    assert rel_maxabs(dtdel, num) < 1e-9


def test_times_equal_with_and_without_gradients():
    # test_transit_timing.jl:91-97
    el, t0 = transit_case()
    h, tmax = 0.04, 2000.0
    _, tt_g, _, _, cg, _ = _run(el, t0, h, tmax, grad=True)
    _, tt_n, _, _, cn, _ = _run(el, t0, h, tmax, grad=False)
    assert np.array_equal(cg, cn)
    assert np.array_equal(tt_g, tt_n)


def test_findtransit_recovers_input_times():
    # test_findtransit.jl:1-25: TRAPPIST-1 (n = 7), h = 0.01, 10 days
    n, t0, h, tmax = 7, 7257.0, 0.01, 10.0
    el = elements()[:n]
    _, tt, _, _, count, _ = _run(el, t0, h, tmax, grad=False)
    for i in range(1, n):
        assert count[i] >= 1
        assert abs((el[i, 2] - tt[i, 0]) / el[i, 2]) < 1e-6


def test_two_body_transits_are_periodic():
    el = elements()[:2].copy()
    el[1, 5] = np.pi / 2   # edge-on: the input t0 is a transit time
    el[1, 6] = 0.0
    P, t0_transit = el[1, 1], el[1, 2]
    t0 = t0_transit - 0.3
    _, tt, dtdq0, dtdel, count, ntt = _run(el, t0, 0.02, 40 * P)
    k = int(count[1])
    assert k >= 39
    times = tt[1, :k]
    assert abs(times[0] - t0_transit) < 1e-10 * abs(t0_transit)
    assert np.max(np.abs(np.diff(times) - P)) < 1e-9
    # d t_k / d t0 = 1 and d t_k / d P = k exactly for a two-body orbit
    assert np.max(np.abs(dtdel[1, :k, 1, 1] - 1.0)) < 1e-8
    assert np.max(np.abs(dtdel[1, :k, 0, 1] - np.arange(k))) < 1e-7 * k


def test_transit_parameters_times_match_transit_timing():
    # TransitParameters (ntbv = 3) shares the search; its first row is tt.
    el, t0 = transit_case()
    h, tmax = 0.04, 30.0
    _, tt1, dq1, de1, c1, _ = _run(el, t0, h, tmax, ntbv=1)
    _, tt3, dq3, de3, c3, _ = _run(el, t0, h, tmax, ntbv=3)
    assert np.array_equal(c1, c3)
    assert np.array_equal(tt1, tt3[0])
    assert np.array_equal(dq1, dq3[0])
    assert np.array_equal(de1, de3[0])
    # impact parameter^2 and sky velocity are positive where transits exist
    for i in range(3):
        k = int(c3[i])
        if k:
            assert np.all(tt3[1, i, :k] > 0) and np.all(tt3[2, i, :k] >= 0)


def test_grad_free_run_keeps_dqdt_and_jacobian():
    """(intr)(s, tt; grad=false) leaves s.dqdt, s.jac_step and s.jac_error as
    they were: ahl21!(s, h) never writes them (ahl21_no_grad.jl:6-18) and
    detect_transits! restores the whole state after each search (set_state!,
    timing.jl:25-26, Integrator.jl:123-134), although findtransit! steps with
    the Jacobian (timing.jl:57)."""
    el, t0 = transit_case()
    x, v, jac_init = O.init_nbody(t0, el)
    s = O.OState(x, v, el[:, 0], t=t0, jac_init=jac_init)
    n7 = 7 * el.shape[0]
    dq_in = np.linspace(-1.0, 1.0, n7)
    s.dqdt[:] = dq_in
    ntt = O.ntt_rule(120.0, el)
    _, _, _, count = O.integrate_transits(s, 0.04, 120.0, ntt, grad=False)
    assert count.sum() > 0
    assert np.array_equal(s.dqdt, dq_in)
    assert np.array_equal(s.jac_step, np.eye(n7)) and np.array_equal(s.jac_error, np.zeros((n7, n7)))


def test_c1_golden_fixture_is_a_transit_record():
    """The committed C1 oracle run (tests/golden/c1_oracle.npz) is what its
    generator says: each planet transits about every period over 4000 days,
    and consecutive times are spaced by its period within 2 %."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_oracle.npz"))
    from nbodygradient_amd import workloads as W
    spec = W.c1()
    P = spec.elements[:, 1]
    for i in range(1, spec.n):
        c = int(g["count"][i])
        assert abs(c - 4000.0 / P[i]) <= 2
        d = np.diff(g["tt"][i, :c])
        assert np.all(np.abs(d / P[i] - 1.0) < 0.02)
    assert g["dtdelements_sel"].shape == (8, (spec.ntt() + 24) // 25, 7, 8)
