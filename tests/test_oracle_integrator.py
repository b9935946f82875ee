"""Pin the oracle's full AHL21 step (ahl21.jl:5-95) and plain integrator.

Restates test/test_integrator.jl (finite differences of the Jacobian after 100
steps and of dq/dt of one step; bitwise equality of grad / no-grad states) and
adds invariants the mathematics fixes: symplecticity of the x-v block, time
reversibility of the palindromic step, and exact two-body motion (phisalpha's
correction vanishes for n = 2, so AHL21 reproduces the Kepler orbit).
"""
import numpy as np
import pytest

import oracle_lib as O
from refcases import integrator_case, rel_maxabs, elements


def test_jacobian_100_steps_vs_finite_differences():
    el, t0, x, v, _ = integrator_case()
    m = el[:, 0].copy()
    h, nstep = 0.05, 100
    s = O.OState(x, v, m, t=t0)
    O.integrate_steps(s, h, nstep, grad=True)
    jac_num = O.fd_steps(x, v, m, h, nstep, dlnq=1e-12)
    # test_integrator.jl:174 compares asinh of both at rtol = sqrt(eps)
    assert rel_maxabs(np.arcsinh(s.jac_step), np.arcsinh(jac_num)) < 1e-11
    assert rel_maxabs(s.jac_step, jac_num) < 1e-11


def test_dqdt_one_step_vs_finite_differences():
    el, t0, x, v, _ = integrator_case()
    m = el[:, 0].copy()
    h = 0.05
    s = O.OState(x, v, m, t=t0)
    O.integrate_steps(s, h, 1, grad=True)
    dq_num = O.fd_step_dqdt(x, v, m, h, dlnq=1e-12)
    assert rel_maxabs(s.dqdt, dq_num) < 1e-11


def test_grad_and_nograd_states_are_bitwise_equal():
    # test_integrator.jl:179-185 (Integrator(h, tmax)(s) from t = t0, 2000 d)
    el = elements()[:3].copy()
    el[1, 0] *= 100.0
    el[2, 0] *= 100.0
    el[:, -1] = 0.0
    t0, h, tmax = 7257.93115525, 0.05, 2000.0
    x, v, _ = O.init_nbody(t0, el)
    m = el[:, 0]
    sg = O.OState(x, v, m, t=t0)
    sn = O.OState(x, v, m, t=t0)
    O.integrate_to(sg, h, t0 + tmax, grad=True)
    O.integrate_to(sn, h, t0 + tmax, grad=False)
    assert np.array_equal(sg.x, sn.x)
    assert np.array_equal(sg.v, sn.v)
    assert sg.t == sn.t == t0 + tmax


def _omega_m(m):
    n = len(m)
    W = np.zeros((6 * n, 6 * n))
    for b in range(n):
        for k in range(3):
            W[6 * b + k, 6 * b + 3 + k] = m[b]
            W[6 * b + 3 + k, 6 * b + k] = -m[b]
    return W


def test_xv_block_is_symplectic():
    el, t0, x, v, _ = integrator_case()
    m = el[:, 0].copy()
    s = O.OState(x, v, m, t=t0)
    O.integrate_steps(s, 0.05, 200, grad=True)
    n = 3
    idx = [7 * b + k for b in range(n) for k in range(6)]
    J = s.jac_step[np.ix_(idx, idx)]
    W = _omega_m(m)
    R = J.T @ W @ J - W
    assert np.max(np.abs(R)) / (np.max(np.abs(J)) ** 2 * np.max(m)) < 1e-12
    # determinant of a symplectic map is one
    assert abs(np.linalg.det(J) - 1.0) < 1e-9


def test_time_reversal():
    # (intr)(s, N) followed by (intr)(s, -N) (Integrator.jl:211-234) returns
    # to the start: AHL21 is a symmetric composition.
    el, t0, x, v, _ = integrator_case()
    m = el[:, 0].copy()
    s = O.OState(x, v, m, t=t0)
    O.integrate_steps(s, 0.05, 400, grad=True)
    O.integrate_steps(s, 0.05, -400, grad=True)
    assert rel_maxabs(s.x, x) < 1e-12
    assert rel_maxabs(s.v, v) < 1e-12
    # and the Jacobian of the round trip is the identity
    assert rel_maxabs(s.jac_step, np.eye(21)) < 1e-10


def test_two_body_motion_is_exact_kepler():
    # n = 2: the drift/kick corrections cancel (ahl21.jl:626-628 with
    # a_ij . r_ij = -G M / r), so the numerical orbit is the Kepler orbit
    # sampled at t0 + k h, whatever the step size.
    el = elements()[:2].copy()
    el[1, 5] = np.pi / 2 + 0.1   # inclined, so all three components are exercised
    el[1, 3] = 0.08               # eccentric
    t0 = 7257.93115525
    x, v, _ = O.init_nbody(t0, el)
    m = el[:, 0].copy()
    for h, nstep in [(0.25, 400), (0.7, 150)]:
        s = O.OState(x, v, m, t=t0)
        O.integrate_steps(s, h, nstep, grad=False)
        rel_x = s.x[:, 1] - s.x[:, 0]
        rel_v = s.v[:, 1] - s.v[:, 0]
        r_k, v_k, _ = O.kepler_init(t0 + nstep * h, m[0] + m[1], el[1, 1:])
        assert np.max(np.abs(rel_x - r_k)) / np.linalg.norm(r_k) < 1e-11
        assert np.max(np.abs(rel_v - v_k)) / np.linalg.norm(v_k) < 1e-11


def test_integrate_to_time_takes_partial_last_step():
    # Integrator.jl:159-197: nsteps = round((time - t0)/h); one extra partial
    # step when t0 + h*nsteps != time.
    el, t0, x, v, _ = integrator_case()
    m = el[:, 0].copy()
    h = 0.05
    s1 = O.OState(x, v, m, t=t0)
    O.integrate_to(s1, h, t0 + 1.02, grad=True)
    s2 = O.OState(x, v, m, t=t0)
    O.integrate_steps(s2, h, 20, grad=True)
    tmax = t0 + h * 20
    O.integrate_steps(s2, (t0 + 1.02) - tmax, 1, grad=True)
    assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.jac_step, s2.jac_step)
    assert s1.t == t0 + 1.02
