"""The workloads C0, C1 and C3 as GPU parity cases.

C0: the single 3-body system, full size (4000 steps), bit for bit against the
    oracle.
C1: the 8-body near-resonant system: bit for bit against the oracle over its
    first 120 days (2000 steps); over its first 1000 days (16,667 steps) the
    transit times with and without gradients are equal bit for bit, as in the
    reference's own tests (test_transit_timing.jl:97).
C3: fit-style vectors around one 6-body base (star + 5 planets): a batch is bit for bit the
    oracle, and the element gradient predicts the change of every transit time
    between two nearby parameter vectors to first order.
(C2 is the batched bench workload, test_gpu_batch.py; C4 the serving tests,
test_gpu_server.py.)
"""
import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu

# our element columns (m, P, t0, ecosw, esinw, I, Omega) -> the element index of
# dtdelements[i, j, l, k] (P, t0, ecosw, esinw, I, Omega, m), init_nbody.jl:146-159
COL_TO_L = [6, 0, 1, 2, 3, 4, 5]


def _gpu_transits(spec, grad=True):
    ic = nb.ElementsIC(spec.t0, spec.n, spec.elements)
    s = nb.State(ic)
    tt = nb.TransitTiming(spec.tmax, ic)
    nb.Integrator(spec.h, spec.t0, spec.tmax)(s, tt, grad=grad)
    return tt, s


def _oracle_transits(spec):
    x, v, jinit = W.initial_states([spec], processes=1)[0]
    os_ = O.OState(x, v, spec.elements[:, 0], t=spec.t0, jac_init=jinit)
    tt, dq, de, count = O.integrate_transits(os_, spec.h, spec.tmax, spec.ntt(), grad=True)
    return tt, dq, de, count, os_


def test_c0_full_size_matches_oracle():
    print("gpu: C0 full size", flush=True)
    spec = W.c0()
    assert spec.nsteps == 4000
    tt, s = _gpu_transits(spec)
    ott, odq, ode, oc, os_ = _oracle_transits(spec)
    assert np.array_equal(tt.count, oc) and oc.sum() > 100
    assert np.array_equal(tt.tt, ott)
    assert np.array_equal(tt.dtdq0, odq)
    assert np.array_equal(tt.dtdelements, ode)
    assert np.array_equal(s.x, os_.x) and np.array_equal(s.v, os_.v)


def test_c1_matches_oracle_then_grad_free_times_agree():
    print("gpu: C1", flush=True)
    spec = W.c1()
    short = W.SystemSpec(spec.elements, spec.t0, spec.h, 120.0)
    tt, s = _gpu_transits(short)
    ott, odq, ode, oc, os_ = _oracle_transits(short)
    assert np.array_equal(tt.count, oc) and oc.sum() > 50
    assert np.array_equal(tt.tt, ott)
    assert np.array_equal(tt.dtdelements, ode)
    longer = W.SystemSpec(spec.elements, spec.t0, spec.h, 1000.0)
    tg, _ = _gpu_transits(longer, grad=True)
    tn, _ = _gpu_transits(longer, grad=False)
    assert np.array_equal(tg.count, tn.count) and tg.count.sum() > 500
    assert np.array_equal(tg.tt, tn.tt)


def test_c3_batch_matches_oracle_and_gradient_predicts_perturbations():
    print("gpu: C3", flush=True)
    specs = W.c3(count=6)
    for sp in specs:
        sp.tmax = 400.0
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    b.run()
    res = b.download()
    b.close()
    for sp, st, r in zip(specs[:3], states[:3], res[:3]):
        x, v, jinit = st
        os_ = O.OState(x, v, sp.elements[:, 0], t=sp.t0, jac_init=jinit)
        ott, odq, ode, oc = O.integrate_transits(os_, sp.h, sp.tmax, sp.ntt(), grad=True)
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["dtdelements"], ode)
    # first-order prediction of a small parameter step from vector 0
    base = specs[0]
    rng = np.random.default_rng(7)
    el1 = base.elements.copy()
    rel = np.zeros_like(el1)
    rel[1:, 0] = 1e-7  # masses
    rel[1:, 1] = 1e-9  # periods
    rel[1:, 2] = 1e-9  # t0
    el1[1:, :3] *= 1.0 + rel[1:, :3] * rng.standard_normal((base.n - 1, 3))
    el1[1:, 3:5] += 1e-8 * rng.standard_normal((base.n - 1, 2))
    d_el = el1 - base.elements
    pert = W.SystemSpec(el1, base.t0, base.h, base.tmax)
    tt1, _ = _gpu_transits(pert)
    r0 = res[0]
    assert np.array_equal(tt1.count, r0["count"])
    de = r0["dtdelements"]  # n x ntt x 7 x n
    pred = np.zeros_like(r0["tt"])
    for k in range(base.n):
        for c in range(7):
            pred += de[:, :, COL_TO_L[c], k] * d_el[k, c]
    dtt = tt1.tt - r0["tt"]
    for i in range(base.n):
        c = int(r0["count"][i])
        if c == 0:
            continue
        scale = np.max(np.abs(dtt[i, :c]))
        assert scale > 0.0
        assert np.max(np.abs(dtt[i, :c] - pred[i, :c])) <= 2e-3 * scale + 1e-11
