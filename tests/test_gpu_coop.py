"""Two-wave (cooperative) workgroups: a system whose transit searches run on a
second wave, from snapshots, gives the same bits as the one-wave kernel and
as the CPU oracle -- alone, for every body count 2..10, and mixed with
one-wave systems in the same batch."""
import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu


def _specs(ns, seed, steps):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        s = W.c2_system(rng, n)
        s.tmax = min(s.tmax, steps * s.h)
        out.append(s)
    return out


def _run(specs, states, mode, max_systems):
    ctx = nb.default_context()
    ctx.set_coop(mode, max_systems)
    try:
        b = Batch(specs, states)
        st = b.run()
        res = b.download(derivatives=True)
        b.close()
    finally:
        ctx.set_coop(1, 128)
    return st, res


def _same(r1, r2):
    for k in ("count", "tt", "dtdq0", "dtdelements", "x", "v"):
        assert np.array_equal(r1[k], r2[k]), k


def test_coop_every_body_count_matches_oracle():
    print("gpu: coop n=2..10 vs oracle", flush=True)
    specs = _specs(range(2, 11), seed=41, steps=500)
    states = W.initial_states(specs, processes=1)
    st, res = _run(specs, states, 2, len(specs))
    assert st["launches"][1]["grid"] == len(specs)  # every system ran on a two-wave workgroup
    for s, (x, v, jinit), r in zip(specs, states, res):
        os_ = O.OState(x, v, s.elements[:, 0], t=s.t0, jac_init=jinit)
        ott, odq, odel, oc = O.integrate_transits(os_, s.h, s.tmax, s.ntt(), grad=True)
        assert oc.sum() > 0
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["dtdq0"], odq)
        assert np.array_equal(r["dtdelements"], odel)
        assert np.array_equal(r["x"], os_.x) and np.array_equal(r["v"], os_.v)


def test_coop_mixed_with_one_wave_systems_equals_coop_off():
    print("gpu: coop mixed batch vs coop off", flush=True)
    specs = _specs([3 + q % 7 for q in range(20)], seed=42, steps=700)
    states = W.initial_states(specs, processes=1)
    _, off = _run(specs, states, 0, 0)
    st, on = _run(specs, states, 2, 5)
    assert st["launches"][1]["grid"] == 5
    for r1, r2 in zip(off, on):
        _same(r1, r2)
    assert st["transits"] == sum(int(np.minimum(r["count"], s.ntt()).sum()) for s, r in zip(specs, on))


def test_coop_long_runs_with_many_transits():
    """Longer runs of the largest body counts: many searches pass through the
    ring (and its four slots are reused many times over)."""
    print("gpu: coop long runs", flush=True)
    specs = _specs([9, 8, 9], seed=43, steps=3000)
    states = W.initial_states(specs, processes=1)
    _, off = _run(specs, states, 0, 0)
    _, on = _run(specs, states, 2, 3)
    for r1, r2 in zip(off, on):
        assert r1["count"].sum() > 50
        _same(r1, r2)


def test_coop_serving_rounds_match_the_oracle():
    """Forced two-wave workgroups in chunked serving rounds: each round ends
    with the searches of its visit finished, and a request resumes (with its
    step cursor, transit counts and gsave) on any workgroup in the next."""
    print("gpu: coop serving rounds vs oracle", flush=True)
    from nbodygradient_amd.server import Server
    specs = _specs([2 + q % 9 for q in range(12)], seed=44, steps=300)
    states = W.initial_states(specs, processes=1)
    ctx = nb.default_context()
    ctx.set_coop(2, 128)
    try:
        sv = Server(max_bytes=1 << 30, chunk=37)
        ids = [sv.submit(s, st) for s, st in zip(specs, states)]
        out = {}
        for _ in range(10000):
            sv.step()
            for rid in sv.finished():
                out[rid] = sv.fetch(rid)
            st = sv.status()
            if st["pending"] == 0 and st["active"] == 0:
                break
        sv.close()
    finally:
        ctx.set_coop(1, 128)
    assert sorted(out) == sorted(ids)
    for rid, s, (x, v, jinit) in zip(ids, specs, states):
        os_ = O.OState(x, v, s.elements[:, 0], t=s.t0, jac_init=jinit)
        ott, odq, odel, oc = O.integrate_transits(os_, s.h, s.tmax, s.ntt(), grad=True)
        r = out[rid]
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["dtdq0"], odq)
        assert np.array_equal(r["dtdelements"], odel)
        assert np.array_equal(r["x"], os_.x)


def test_coop_wait_bound_gives_an_error_not_a_hang():
    """With every inter-wave wait bounded to a single poll, the searching wave
    gives up before the first transit and the integrating wave then finds the
    ring full: the kernel still ends, and the run reports the failure.  The
    next run, with the default bounds, is the usual one."""
    print("gpu: coop wait bound", flush=True)
    specs = _specs([9, 8], seed=45, steps=600)
    states = W.initial_states(specs, processes=1)
    ctx = nb.default_context()
    ctx.set_coop(2, 2)
    ctx.set_coop_wait(1)
    try:
        b = Batch(specs, states)
        with pytest.raises(nb.NbgError) as ei:
            b.run()
        assert "timed out" in str(ei.value)
        b.close()
    finally:
        ctx.set_coop_wait(0)
        ctx.set_coop(1, 128)
    _, off = _run(specs, states, 0, 0)
    _, on = _run(specs, states, 2, 2)
    for r1, r2 in zip(off, on):
        assert r1["count"].sum() > 4  # enough transits to fill the ring
        _same(r1, r2)
