"""The oracle's CartesianOutput run (Outputs.jl:26-49), pinned without the
reference (which has no test of it): record 0 is the initial state; record
k+1 is exactly one oracle step from record k (so each record holds the whole
state before its step, Kahan errors included); t_k = t0 + h*k; the final
state equals N plain steps; the direction follows check_step(t0, tmax); and
the no-grad run gives the same x and v (test_integrator.jl:184-185)."""
import numpy as np

import oracle_lib as O
from nbodygradient_amd import workloads as W


def _system(n, seed):
    rng = np.random.default_rng(seed)
    sp = W.c2_system(rng, n)
    x, v, _ = W.initial_states([sp], processes=1)[0]
    return sp, x, v


def _from_record(rec, k, m):
    s = O.OState(rec["x"][:, :, k], rec["v"][:, :, k], m, t=rec["t"][k], jac=rec["jac_step"][:, :, k])
    s.xerror = np.asfortranarray(rec["xerror"][:, :, k].copy())
    s.verror = np.asfortranarray(rec["verror"][:, :, k].copy())
    s.jac_error = np.asfortranarray(rec["jac_error"][:, :, k].copy())
    return s


def test_records_are_the_state_before_each_step():
    sp, x, v = _system(4, 11)
    m = sp.elements[:, 0]
    s = O.OState(x, v, m, t=sp.t0)
    N = 40
    rec = O.cartesian_output(s, sp.h, sp.t0 + 100.0, N)
    assert np.array_equal(rec["x"][:, :, 0], x) and np.array_equal(rec["v"][:, :, 0], v)
    assert np.array_equal(rec["jac_step"][:, :, 0], np.eye(28)) and not rec["jac_error"][:, :, 0].any()
    assert rec["t"][0] == sp.t0
    assert np.array_equal(rec["t"], sp.t0 + sp.h * np.arange(N, dtype=np.float64))
    assert s.t == sp.t0 + sp.h * N
    for k in range(N - 1):
        c = _from_record(rec, k, m)
        O.step(c, sp.h)
        assert np.array_equal(c.x, rec["x"][:, :, k + 1]) and np.array_equal(c.v, rec["v"][:, :, k + 1])
        assert np.array_equal(c.xerror, rec["xerror"][:, :, k + 1])
        assert np.array_equal(c.jac_step, rec["jac_step"][:, :, k + 1])
        assert np.array_equal(c.jac_error, rec["jac_error"][:, :, k + 1])
    # the state after the run is N plain steps
    p = O.OState(x, v, m, t=sp.t0)
    O.integrate_steps(p, sp.h, N)
    assert np.array_equal(p.x, s.x) and np.array_equal(p.jac_step, s.jac_step)
    c = _from_record(rec, N - 1, m)
    O.step(c, sp.h)
    assert np.array_equal(c.x, s.x) and np.array_equal(c.jac_error, s.jac_error)


def test_direction_and_no_grad():
    sp, x, v = _system(3, 12)
    m = sp.elements[:, 0]
    # |tmax| > |t0| with tmax < 0: backward steps (check_step = sign(tmax))
    s = O.OState(x, v, m, t=1.0)
    rec = O.cartesian_output(s, sp.h, -50.0, 25)
    assert rec["t"][1] == 1.0 - sp.h and s.t == 1.0 + (-sp.h * 25)
    b = O.OState(x, v, m, t=1.0)
    O.integrate_steps(b, sp.h, -25)
    assert np.array_equal(b.x, s.x) and np.array_equal(b.v, s.v)
    # same sign, |tmax| < |t0|: also backward (check_step = -sign(tmax))
    s2 = O.OState(x, v, m, t=10.0)
    rec2 = O.cartesian_output(s2, sp.h, 5.0, 3)
    assert rec2["t"][2] == 10.0 + (-sp.h * 2)
    g = O.OState(x, v, m, t=sp.t0)
    ng = O.OState(x, v, m, t=sp.t0)
    rg = O.cartesian_output(g, sp.h, sp.t0 + 10.0, 30, grad=True)
    rn = O.cartesian_output(ng, sp.h, sp.t0 + 10.0, 30, grad=False, jacobian=False)
    assert np.array_equal(rg["x"], rn["x"]) and np.array_equal(rg["v"], rn["v"])
    assert np.array_equal(g.x, ng.x)
