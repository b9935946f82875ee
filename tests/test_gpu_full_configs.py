"""The workloads at their full sizes on the GPU, against the oracle.

C1: the 8-body near-resonant system over its whole run (66,667 steps of
    h = 0.06 d to 4000 d, 7,557 transits), one system per launch, against the
    oracle's run of the same system (tests/golden/c1_oracle.npz, written by
    tests/golden/make_c1_golden.py): counts, every transit time and the
    dtdelements rows of every 25th transit.
C3: fit-batch vectors (perturbations of the 6-body base: star + 5 planets)
    at the full tmax = 3000 d (30,000 steps each) as a batch, against the
    oracle run of the same vectors.
C4: serving-trace requests (n ~ U{2..10}, tmax log-uniform in [100, 8000] d)
    at their full tmax through the server, against the oracle.
Everything is compared bit for bit; the tolerances the reference is held to
(transit times 1e-12 relative, gradients 1e-10 relative of each array's
largest entry) are checked too, so that a failure says which bar it missed.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch
from nbodygradient_amd.server import Server

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _tol(a, b, rel):
    """max |a - b| <= rel * max |b| (the reference's maxabs norm)."""
    return np.max(np.abs(a - b)) <= rel * np.max(np.abs(b))


def _oracle(spec, state):
    x, v, jinit = state
    os_ = O.OState(x, v, spec.elements[:, 0], t=spec.t0, jac_init=jinit)
    tt, dq, de, count = O.integrate_transits(os_, spec.h, spec.tmax, spec.ntt(), grad=True)
    return {"tt": tt, "dtdelements": de, "count": count, "x": os_.x, "v": os_.v}


def test_c1_full_run_matches_oracle():
    print("gpu: C1 full run (66,667 steps)", flush=True)
    g = np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))
    spec = W.c1()
    assert spec.nsteps == int(g["nsteps"])
    ic = nb.ElementsIC(spec.t0, spec.n, spec.elements)
    s = nb.State(ic)
    tt = nb.TransitTiming(spec.tmax, ic)
    nb.Integrator(spec.h, spec.t0, spec.tmax)(s, tt, grad=True)
    assert np.array_equal(tt.count, g["count"]) and tt.count.sum() > 7000
    assert _tol(tt.tt, g["tt"], 1e-12)
    assert np.array_equal(tt.tt, g["tt"])
    sel = tt.dtdelements[:, ::int(g["every"])]
    assert _tol(sel, g["dtdelements_sel"], 1e-10)
    assert np.array_equal(sel, g["dtdelements_sel"])
    assert np.array_equal(s.x, g["x"]) and np.array_equal(s.v, g["v"])


def test_c3_full_tmax_batch_matches_oracle():
    print("gpu: C3 vectors at tmax 3000", flush=True)
    specs = W.c3(16384)[:4]
    assert all(sp.nsteps == 30000 for sp in specs)
    states = W.initial_states(specs, processes=1)
    with ThreadPoolExecutor(4) as ex:  # the oracle runs on the host while the GPU works
        fut = [ex.submit(_oracle, sp, st) for sp, st in zip(specs, states)]
        b = Batch(specs, states)
        b.run()
        res = b.download()
        b.close()
        ref = [f.result() for f in fut]
    for r, o in zip(res, ref):
        assert np.array_equal(r["count"], o["count"]) and o["count"].sum() > 1000
        assert _tol(r["tt"], o["tt"], 1e-12) and _tol(r["dtdelements"], o["dtdelements"], 1e-10)
        assert np.array_equal(r["tt"], o["tt"])
        assert np.array_equal(r["dtdelements"], o["dtdelements"])
        assert np.array_equal(r["x"], o["x"]) and np.array_equal(r["v"], o["v"])


def test_c4_full_tmax_requests_through_the_server_match_oracle():
    print("gpu: C4 requests at full tmax", flush=True)
    trace = W.c4(count=400)
    # a handful of requests of different body counts and lengths (up to
    # 20,000 steps, so that the oracle finishes in seconds)
    picked, seen = [], set()
    for _, sp in trace:
        if sp.n not in seen and 3000 <= sp.nsteps <= 20000:
            picked.append(sp)
            seen.add(sp.n)
        if len(picked) == 5:
            break
    assert len(picked) == 5
    states = W.initial_states(picked, processes=1)
    with ThreadPoolExecutor(5) as ex:
        fut = [ex.submit(_oracle, sp, st) for sp, st in zip(picked, states)]
        sv = Server(max_bytes=1 << 30, chunk=256)
        ids = [sv.submit(sp, st) for sp, st in zip(picked, states)]
        out = {}
        for _ in range(100000):
            sv.step()
            for rid in sv.finished():
                out[rid] = sv.fetch(rid)
            if len(out) == len(ids):
                break
        sv.close()
        ref = [f.result() for f in fut]
    for rid, o in zip(ids, ref):
        r = out[rid]
        assert np.array_equal(r["count"], o["count"])
        assert np.array_equal(r["tt"], o["tt"])
        assert np.array_equal(r["dtdelements"], o["dtdelements"])
        assert np.array_equal(r["x"], o["x"]) and np.array_equal(r["v"], o["v"])
