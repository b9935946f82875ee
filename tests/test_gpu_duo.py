"""Two systems per wave (nbg_set_systems_per_wave(2), the nbg_run_duo
launch): every output of a batch -- transit times and counts, dt/dq0,
dt/delements, the final x and v -- is the same bits as with one system per
wave, which test_gpu_batch.py compares with the oracle (and a few paired
systems are compared with the oracle here directly).  The batches cover the
paired body counts (n <= 9, whose Jacobians then live in HBM) beside unpaired
ones (n = 10..12, and the n = 10 launch's own systems), odd counts of a body
count, different step sizes and run lengths within a pair, kicked pairs with
the same and with different pair masks, massless bodies (sequential sweep
order: no pairing), TransitParameters (ntbv = 3), runs without gradients,
rounds with restarts, and the autotuner choosing between the two."""
import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu

KEYS = ("tt", "count", "dtdq0", "dtdelements", "x", "v")


def _specs(ns, seed, steps_lo=60, steps_hi=420):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        s = W.c2_system(rng, n)
        s.tmax = int(rng.integers(steps_lo, steps_hi)) * s.h
        out.append(s)
    return out


# Pairing keeps at least one unit per wave slot and pairs only runs short
# against the batch's work per slot, so batches this small pair systems only
# on a small grid: every run here (both forms) uses `grid` workgroups.
GRID = 2


def _run(specs, k, rounds=None, derivatives=True, states=None, grid=GRID, **kw):
    ctx = nb.default_context()
    ctx.set_systems_per_wave(k)
    ctx.set_grid_cap(grid)
    try:
        b = Batch(specs, states if states is not None else W.initial_states(specs, processes=1), **kw)
        if rounds is None:
            st = b.run()
        else:
            st = None
            for K, restart in rounds:
                s_ = b.round(K, restart=restart)
                st = s_ if st is None else st
        res = b.download(derivatives=derivatives)
        b.close()
    finally:
        ctx.set_systems_per_wave(1)
        ctx.set_grid_cap(0)
    if k == 2:
        assert st["paired_systems"] > 0
    else:
        assert st["paired_systems"] == 0
    return st, res


def _same(ra, rb, keys=KEYS):
    assert len(ra) == len(rb)
    for a, c in zip(ra, rb):
        for k in keys:
            if k in a:
                assert np.array_equal(a[k], c[k]), k


MIX = [9, 9, 9, 8, 8, 7, 7, 7, 2, 2, 2, 3, 4, 5, 6, 9, 8, 6, 5, 3, 10, 10, 10, 11, 12]


@pytest.mark.parametrize("grad", [True, False])
def test_duo_equals_one_system_per_wave(grad):
    print("gpu: duo vs solo grad=%s" % grad, flush=True)
    specs = _specs(MIX, seed=41)
    st1, r1 = _run(specs, 1, grad=grad, derivatives=grad)
    st2, r2 = _run(specs, 2, grad=grad, derivatives=grad)
    _same(r1, r2)
    assert st1["transits"] == st2["transits"] and st1["newton_iters"] == st2["newton_iters"]
    assert sum(int(r["count"].sum()) for r in r1) > 20


def test_duo_paired_systems_match_oracle():
    print("gpu: duo vs oracle", flush=True)
    specs = _specs([9, 9, 7, 7, 2, 2], seed=42, steps_lo=150, steps_hi=300)
    states = W.initial_states(specs, processes=1)
    _, res = _run(specs, 2, grid=1)
    for s, (x, v, jinit), r in zip(specs, states, res):
        os_ = O.OState(x, v, s.elements[:, 0], t=s.t0, jac_init=jinit)
        ott, odq, odel, oc = O.integrate_transits(os_, s.h, s.tmax, s.ntt(), grad=True)
        assert np.array_equal(r["count"], oc)
        assert np.array_equal(r["tt"], ott)
        assert np.array_equal(r["dtdq0"], odq)
        assert np.array_equal(r["dtdelements"], odel)
        assert np.array_equal(r["x"], os_.x) and np.array_equal(r["v"], os_.v)


def test_duo_kicked_pairs_and_massless_bodies():
    """Pair masks decide the level schedule: equal masks pair, different
    ones (or two massless bodies, whose pair keeps the sequential order) run
    one after the other; the results do not depend on it."""
    print("gpu: duo kicked / massless", flush=True)
    specs = _specs([8, 8, 8, 8, 7, 7, 9, 9], seed=43)
    pairs = []
    for q, s in enumerate(specs):
        n = s.n
        p = np.zeros((n, n), dtype=np.uint8)
        if q in (0, 1):  # the same kicked planet pair in both
            p[2, 3] = p[3, 2] = 1
        if q == 2:  # another
            p[1, 4] = p[4, 1] = 1
        pairs.append(p)
    states = W.initial_states(specs, processes=1)
    el = np.array(specs[6].elements, dtype=float)
    el[2, 0] = el[3, 0] = 0.0  # two massless planets (initial state from the massive elements)
    specs[6] = W.SystemSpec(el, specs[6].t0, specs[6].h, specs[6].tmax)
    _, r1 = _run(specs, 1, pair=pairs, states=states)
    _, r2 = _run(specs, 2, pair=pairs, states=states)
    for r in r1:
        assert np.all(np.isfinite(r["x"])) and np.all(np.isfinite(r["tt"]))
    _same(r1, r2)


def test_duo_transit_parameters():
    print("gpu: duo ntbv=3", flush=True)
    specs = _specs([9, 9, 8, 8, 7, 2, 2, 5], seed=44)
    _, r1 = _run(specs, 1, ntbv=3)
    _, r2 = _run(specs, 2, ntbv=3)
    _same(r1, r2)


def test_duo_rounds_with_restart():
    """Rounds: pairs advance in lockstep until one run ends; restarted
    systems begin again from their initial records."""
    print("gpu: duo rounds", flush=True)
    specs = _specs(MIX, seed=45, steps_lo=60, steps_hi=400)
    plan = [(150, True)] * 4 + [(150, False)] * 4
    _, r1 = _run(specs, 1, rounds=plan)
    _, r2 = _run(specs, 2, rounds=plan)
    _same(r1, r2)


def test_autotune_round_tries_two_systems_per_wave():
    print("gpu: autotune with duo", flush=True)
    specs = _specs(MIX, seed=46)
    states = W.initial_states(specs, processes=1)
    ctx = nb.default_context()
    ctx.set_grid_cap(GRID)
    try:
        b = Batch(specs, states)
        b.run()
        ref = b.download()
        tune = b.autotune_round(100)
        assert {c["systems_per_wave"] for c in tune["candidates"]} == {1, 2}
        assert all(c["ms"] > 0 for c in tune["candidates"])
        b.run()
        _same(b.download(), ref)
        b.close()
    finally:
        ctx.set_coop(1, 128)
        ctx.set_launch(0)
        ctx.set_systems_per_wave(1)
        ctx.set_grid_cap(0)


def test_pairing_keeps_a_unit_per_wave_slot():
    """Without a grid cap a batch smaller than the GPU's wave slots is not
    paired (pairs would leave slots idle); results are the same either way."""
    print("gpu: duo pairing policy", flush=True)
    specs = _specs([5] * 40, seed=47)
    states = W.initial_states(specs, processes=1)
    ctx = nb.default_context()
    ctx.set_systems_per_wave(2)
    try:
        b = Batch(specs, states)
        st = b.run()
        assert st["paired_systems"] == 0
        r2 = b.download()
        b.close()
    finally:
        ctx.set_systems_per_wave(1)
    _, r1 = _run(specs, 1, states=states)
    _same(r1, r2)
