"""Batched form on the GPU: every system of a mixed batch equals the CPU oracle
bit for bit, and equals the same system run alone (plain form)."""
import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu


def _mixed(count=24, seed=11, steps=400):
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(count):
        s = W.c2_system(rng, 3 + q % 7)
        s.tmax = min(s.tmax, steps * s.h)
        specs.append(s)
    return specs


@pytest.mark.parametrize("derivatives", [True, False])
def test_batch_matches_oracle(derivatives):
    print("gpu: batch vs oracle derivatives=%s" % derivatives, flush=True)
    specs = _mixed()
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    st = b.run()
    res = b.download(derivatives=derivatives)
    assert st["transits"] == sum(int(np.minimum(r["count"], s.ntt()).sum()) for s, r in zip(specs, res))
    for s, (x, v, jinit), r in zip(specs, states, res):
        os_ = O.OState(x, v, s.elements[:, 0], t=s.t0, jac_init=jinit)
        ott, odq, odel, oc = O.integrate_transits(os_, s.h, s.tmax, s.ntt(), grad=True)
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["x"], os_.x) and np.array_equal(r["v"], os_.v)
        if derivatives:
            assert np.array_equal(r["dtdq0"], odq)
            assert np.array_equal(r["dtdelements"], odel)
    # a second run after reset reproduces the first
    b.reset()
    b.run()
    res2 = b.download(derivatives=derivatives)
    for r1, r2 in zip(res, res2):
        assert np.array_equal(r1["tt"], r2["tt"]) and np.array_equal(r1["count"], r2["count"])
    b.close()


@pytest.mark.parametrize("n", list(range(2, 13)))
def test_systems_in_sequence_on_one_workgroup(n):
    """One persistent workgroup integrates several systems back to back: every
    register and LDS value carried from one system to the next must be
    re-established (regression: a copy of the LDS address of the transit
    counts was once lost between systems for n = 9)."""
    print("gpu: sequence on one workgroup n=%d" % n, flush=True)
    rng = np.random.default_rng(100 + n)
    specs = [W.c2_system(rng, n) for _ in range(3)]
    for s, k in zip(specs, (140, 4, 140)):
        s.tmax = k * s.h
    states = W.initial_states(specs, processes=1)
    ctx = nb.default_context()
    ctx.set_grid_cap(1)
    try:
        b = Batch(specs, states)
        b.run()
        res = b.download()
        b.close()
    finally:
        ctx.set_grid_cap(0)
    for s, (x, v, jinit), r in zip(specs, states, res):
        os_ = O.OState(x, v, s.elements[:, 0], t=s.t0, jac_init=jinit)
        ott, odq, odel, oc = O.integrate_transits(os_, s.h, s.tmax, s.ntt(), grad=True)
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["x"], os_.x) and np.array_equal(r["v"], os_.v)
        assert np.array_equal(r["dtdq0"], odq)
        assert np.array_equal(r["dtdelements"], odel)


def test_same_alone_and_in_the_mix():
    print("gpu: alone vs mix", flush=True)
    specs = _mixed(count=16, seed=12, steps=300)
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    b.run()
    res = b.download()
    for k in (0, 5, 11):
        s = specs[k]
        ic = nb.ElementsIC(s.t0, s.n, s.elements)
        st = nb.State(ic)
        tt = nb.TransitTiming(s.tmax, ic)
        nb.Integrator(s.h, s.tmax)(st, tt)
        assert np.array_equal(tt.tt, res[k]["tt"])
        assert np.array_equal(tt.dtdelements, res[k]["dtdelements"])
        assert np.array_equal(tt.count, res[k]["count"])
    b.close()
