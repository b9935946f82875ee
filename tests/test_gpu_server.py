"""Serving (continuous batching) on the GPU: every request's result equals the
one-system run bit for bit, alone or in a mix, whatever the round length and
the admission order imposed by a small arena."""
import numpy as np
import pytest

import oracle_lib as O
from nbodygradient_amd import workloads as W
from nbodygradient_amd.server import Server, serve

pytestmark = pytest.mark.gpu


def _requests(count, seed, steps, nmin=2, nmax=10):
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(count):
        s = W.c2_system(rng, nmin + q % (nmax - nmin + 1))
        s.tmax = min(s.tmax, steps * s.h)
        specs.append(s)
    return specs, W.initial_states(specs, processes=1)


def _oracle(spec, state):
    x, v, jinit = state
    os_ = O.OState(x, v, spec.elements[:, 0], t=spec.t0, jac_init=jinit)
    tt, dq, de, count = O.integrate_transits(os_, spec.h, spec.tmax, spec.ntt(), grad=True)
    return tt, dq, de, count, os_.x


def _drain(server):
    out = {}
    for _ in range(100000):
        nf, _ = server.step()
        for rid in server.finished():
            out[rid] = server.fetch(rid)
        st = server.status()
        if st["pending"] == 0 and st["active"] == 0:
            break
    return out


def test_served_requests_match_the_oracle():
    print("gpu: server vs oracle", flush=True)
    specs, states = _requests(18, seed=31, steps=300)
    sv = Server(max_bytes=1 << 30, chunk=37)
    ids = [sv.submit(s, st) for s, st in zip(specs, states)]
    res = _drain(sv)
    assert sorted(res) == sorted(ids)
    for rid, s, st in zip(ids, specs, states):
        tt, dq, de, count, x = _oracle(s, st)
        r = res[rid]
        assert np.array_equal(r["count"], count)
        assert np.array_equal(r["tt"], tt)
        assert np.array_equal(r["dtdq0"], dq)
        assert np.array_equal(r["dtdelements"], de)
        assert np.array_equal(r["x"], x)
    sv.close()


def test_same_alone_and_in_the_mix():
    print("gpu: server alone vs mix", flush=True)
    specs, states = _requests(12, seed=32, steps=260)
    alone = {}
    for k in (0, 4, 9):
        sv = Server(max_bytes=256 << 20, chunk=1000)
        rid = sv.submit(specs[k], states[k])
        alone[k] = _drain(sv)[rid]
        sv.close()
    sv = Server(max_bytes=1 << 30, chunk=23)
    # requests arrive while others run: half now, the rest after two rounds
    ids = {k: sv.submit(specs[k], states[k]) for k in range(6)}
    sv.step()
    sv.step()
    ids.update({k: sv.submit(specs[k], states[k]) for k in range(6, 12)})
    mix = _drain(sv)
    for k in alone:
        a, b = alone[k], mix[ids[k]]
        for key in ("tt", "dtdq0", "dtdelements", "count", "x", "v"):
            assert np.array_equal(a[key], b[key]), key
    sv.close()


def test_admission_by_footprint_and_real_time_driver():
    print("gpu: server admission", flush=True)
    specs, states = _requests(8, seed=33, steps=200, nmin=5, nmax=8)
    big = max(Server.footprint(s.n, s.ntt()) for s in specs)
    sv = Server(max_bytes=2 * big + 4096, chunk=64)  # at most about two requests resident
    arrivals = np.linspace(0.0, 0.05, len(specs))
    out = serve(sv, specs, states, arrivals, keep_results=True)
    assert np.all(np.isfinite(out["latency_s"])) and np.all(out["latency_s"] >= 0.0)
    assert len(out["results"]) == len(specs)
    for k in (0, 3, 7):
        tt, dq, de, count, x = _oracle(specs[k], states[k])
        r = out["results"][k]
        assert np.array_equal(r["tt"], tt) and np.array_equal(r["dtdelements"], de)
        assert np.array_equal(r["count"], count)
    sv.close()
