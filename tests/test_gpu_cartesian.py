"""CartesianOutput on the GPU (nbg_integrate_cartesian): the state before
every step, bit for bit equal to the oracle's run of the same records
(Outputs.jl:26-49), for body counts across every launch class, both
directions, without the Jacobian records, and without grad."""
import sys

import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _system(n, seed):
    rng = np.random.default_rng(seed)
    sp = W.c2_system(rng, n)
    x, v, _ = W.initial_states([sp], processes=1)[0]
    return sp, x, v


def _pstate(x, v, m, t):
    n = x.shape[1]
    s = nb.State.__new__(nb.State)
    s.n = n
    s.x = np.asfortranarray(x.copy())
    s.v = np.asfortranarray(v.copy())
    s.t = np.array([t])
    s.m = np.asarray(m, dtype=np.float64).copy()
    s.jac_step = np.asfortranarray(np.eye(7 * n))
    s.dqdt = np.zeros(7 * n)
    s.jac_init = np.asfortranarray(np.eye(7 * n))
    s.pair = np.zeros((n, n), dtype=bool, order="F")
    s.xerror = np.zeros((3, n), order="F")
    s.verror = np.zeros((3, n), order="F")
    s.dqdt_error = np.zeros(7 * n)
    s.jac_error = np.zeros((7 * n, 7 * n), order="F")
    return s


def _check(o, rec, jac=True):
    assert np.array_equal(o.t, rec["t"])
    for k in ("x", "v", "xerror", "verror"):
        assert np.array_equal(getattr(o, k), rec[k]), k
    if jac:
        assert np.array_equal(o.jac_step, rec["jac_step"])
        assert np.array_equal(o.jac_error, rec["jac_error"])


@pytest.mark.parametrize("n", [2, 3, 6, 7, 9, 10, 12])
def test_cartesian_matches_oracle(n):
    say("gpu: cartesian n=%d" % n)
    sp, x, v = _system(n, 60 + n)
    m = sp.elements[:, 0]
    N = 120
    ps = _pstate(x, v, m, sp.t0)
    o = nb.CartesianOutput(n, N)
    nb.Integrator(sp.h, sp.t0 + 1000.0)(ps, o)
    os_ = O.OState(x, v, m, t=sp.t0)
    rec = O.cartesian_output(os_, sp.h, sp.t0 + 1000.0, N)
    _check(o, rec)
    assert ps.t[0] == os_.t
    assert np.array_equal(ps.x, os_.x) and np.array_equal(ps.jac_step, os_.jac_step)
    assert np.array_equal(ps.jac_error, os_.jac_error) and np.array_equal(ps.verror, os_.verror)
    s0 = o.state(0)
    assert np.array_equal(s0.x, x) and np.array_equal(s0.jac_step, np.eye(7 * n))


def test_cartesian_backward_no_jacobian_and_no_grad():
    say("gpu: cartesian backward / no-jac / no-grad")
    sp, x, v = _system(5, 71)
    m = sp.elements[:, 0]
    N = 80
    # backward: tmax < 0 with |tmax| > |t0|
    ps = _pstate(x, v, m, 2.0)
    o = nb.CartesianOutput(5, N)
    nb.Integrator(sp.h, -500.0)(ps, o)
    os_ = O.OState(x, v, m, t=2.0)
    rec = O.cartesian_output(os_, sp.h, -500.0, N)
    _check(o, rec)
    assert o.t[1] < o.t[0] and ps.t[0] == os_.t
    # records of t, x, v and errors only
    ps = _pstate(x, v, m, sp.t0)
    o2 = nb.CartesianOutput(5, N, jacobian=False)
    nb.Integrator(sp.h, sp.t0 + 100.0)(ps, o2)
    os_ = O.OState(x, v, m, t=sp.t0)
    rec = O.cartesian_output(os_, sp.h, sp.t0 + 100.0, N)
    _check(o2, rec, jac=False)
    assert np.array_equal(ps.jac_step, os_.jac_step)
    # without grad: the same x and v
    ps = _pstate(x, v, m, sp.t0)
    o3 = nb.CartesianOutput(5, N, jacobian=False)
    nb.Integrator(sp.h, sp.t0 + 100.0)(ps, o3, grad=False)
    assert np.array_equal(o3.x, rec["x"]) and np.array_equal(o3.v, rec["v"])
    assert np.array_equal(ps.x, os_.x)
    # nstep = 0 leaves the state as it was
    ps = _pstate(x, v, m, sp.t0)
    nb.Integrator(sp.h, sp.t0 + 100.0)(ps, nb.CartesianOutput(5, 0))
    assert np.array_equal(ps.x, x) and ps.t[0] == sp.t0


def test_cartesian_batch_matches_oracle():
    """Many systems' records in one persistent launch (nbg_batch_cartesian),
    body counts across the launch classes, bitwise equal to the oracle."""
    say("gpu: cartesian batch")
    from nbodygradient_amd.batch import cartesian_batch
    rng = np.random.default_rng(97)
    specs = [W.c2_system(rng, n) for n in (3, 6, 8, 10, 2)]
    nstep = [40, 25, 31, 17, 50]
    states = W.initial_states(specs, processes=1)
    res = cartesian_batch(specs, nstep, states, grad=True, errors=True, jacobian=True)
    for sp, N, (x, v, _), r in zip(specs, nstep, states, res):
        os_ = O.OState(x, v, sp.elements[:, 0], t=sp.t0)
        rec = O.cartesian_output(os_, sp.h, sp.tmax, N)
        for k in ("t", "x", "v", "xerror", "verror", "jac_step", "jac_error"):
            assert np.array_equal(r[k], rec[k]), (sp.n, k)
