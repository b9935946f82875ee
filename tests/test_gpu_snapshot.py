"""TransitSnapshot through the product path (integration to each time on
the GPU, nbg_integrate_to) against the oracle's run of the same snapshot,
bit for bit: b^2, v_sky, dbvdq0 and dbvdelements, with and without grad."""
import os

import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _trappist():
    el = np.loadtxt(os.path.join(ROOT, "tests", "golden", "elements.txt"), delimiter=",")[:3].copy()
    el[1:, 2] -= 7300.0
    el[:, -1] = 0.0
    el[1, 0] *= 10.0
    el[2, 0] *= 10.0
    return el, 7257.93115525 - 7300.0 - 0.5, 0.04


@pytest.mark.parametrize("case", ["trappist", "c2n6"])
@pytest.mark.parametrize("grad", [True, False])
def test_snapshot_matches_oracle(case, grad):
    print("gpu: snapshot %s grad=%s" % (case, grad), flush=True)
    if case == "trappist":
        el, t0, h = _trappist()
        times = np.array([t0 + 0.35, t0 + 1.0, t0 + 4.25, t0 + 4.25, t0 + 9.9])
    else:
        sp = W.c2_system(np.random.default_rng(95), 6)
        el, t0, h = sp.elements[:, :7], sp.t0, sp.h
        times = t0 + np.array([3.1, 17.0, 17.05, 40.2]) * h
    ic = nb.ElementsIC(t0, el.shape[0], el)
    s = nb.State(ic)
    ts = nb.TransitSnapshot(times, ic)
    nb.Integrator(h, 0.0, 10.0)(s, ts, grad=grad)
    # the product's initial conditions come from its own host stage; start the
    # oracle from the same numbers so the comparison is of the integration
    s0 = nb.State(ic)
    os_ = O.OState(s0.x, s0.v, el[:, 0], t=t0, jac_init=s0.jac_init)
    vsky, bsky2, dq, de = O.transit_snapshot(os_, h, times, grad=grad)
    assert np.array_equal(ts.vsky, vsky) and np.array_equal(ts.bsky2, bsky2)
    assert np.array_equal(ts.dbvdq0, dq) and np.array_equal(ts.dbvdelements, de)
    assert s.t[0] == times[-1] and np.array_equal(s.x, os_.x)
    if grad:
        assert np.abs(ts.dbvdq0[0, 1:]).max() > 0


def test_snapshot_batch_matches_oracle():
    """Many systems' snapshots in one persistent launch (nbg_batch_snapshot),
    body counts across the launch classes, each bitwise equal to the oracle."""
    print("gpu: snapshot batch", flush=True)
    from nbodygradient_amd.batch import snapshot_batch
    rng = np.random.default_rng(96)
    specs, times = [], []
    for q, n in enumerate([3, 5, 7, 9, 10, 4, 8]):
        sp = W.c2_system(rng, n)
        specs.append(sp)
        times.append(sp.t0 + np.sort(rng.uniform(0.0, 60.0, size=2 + q % 3)) * sp.h)
    states = W.initial_states(specs, processes=1)
    for grad in (True, False):
        res = snapshot_batch(specs, times, states, grad=grad)
        for sp, tj, (x, v, ji), r in zip(specs, times, states, res):
            os_ = O.OState(x, v, sp.elements[:, 0], t=sp.t0, jac_init=ji)
            vs, bs, dq, de = O.transit_snapshot(os_, sp.h, tj, grad=grad)
            assert np.array_equal(r["vsky"], vs) and np.array_equal(r["bsky2"], bs), sp.n
            if grad:
                assert np.array_equal(r["dbvdq0"], dq) and np.array_equal(r["dbvdelements"], de), sp.n
