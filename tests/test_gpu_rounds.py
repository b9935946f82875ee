"""Rounds of a batch (nbg_batch_round, the benchmark's unit of work): running a
batch in rounds of K steps gives the same bits as running it to completion
(nbg_batch_run, itself compared with the oracle in test_gpu_batch.py), the
rounds' work counts add up to the whole run's, and a system restarted from
its initial record after its run ended repeats that run exactly."""
import numpy as np
import pytest

from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu

KEYS = ("tt", "count", "dtdq0", "dtdelements", "x", "v")


def _specs(count=20, seed=31):
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(count):
        s = W.c2_system(rng, 3 + q % 7)
        s.tmax = (60 + 37 * q) * s.h  # uneven step counts: 60 .. 763
        specs.append(s)
    return specs


def _same(r1, r2):
    for k in KEYS:
        assert np.array_equal(r1[k], r2[k]), k


def test_rounds_equal_whole_run():
    print("gpu: rounds vs whole run", flush=True)
    specs = _specs()
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    whole = b.run()
    ref = b.download()
    b.reset()
    steps = trans = newton = 0
    for r in range(40):
        st = b.round(150, restart=False)
        steps += st["system_steps"]
        trans += st["transits"]
        newton += st["newton_iters"]
        if st["system_steps"] == 0:
            break
    assert steps == whole["system_steps"]
    assert trans == whole["transits"] and newton == whole["newton_iters"]
    for a, c in zip(b.download(), ref):
        _same(a, c)
    b.close()


def test_round_units_by_body_count():
    """One round from the start: every system advances min(K, its steps);
    the per-launch units by n are those steps plus the round's transits."""
    print("gpu: round units", flush=True)
    specs = _specs(count=14, seed=32)
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    K = 200
    st = b.round(K, restart=False)
    by_n = {}
    for L in st["launches"].values():
        for n, u in L["units"].items():
            by_n[n] = by_n.get(n, 0) + u
    main = {}
    for s in specs:
        main[s.n] = main.get(s.n, 0) + min(K, s.nsteps)
    res = b.download(derivatives=False)
    tr = {}
    for s, r in zip(specs, res):
        tr[s.n] = tr.get(s.n, 0) + int(np.minimum(r["count"], s.ntt()).sum())
    assert st["transits"] == sum(tr.values())
    assert by_n == {n: main[n] + tr.get(n, 0) for n in main}
    b.close()


def test_restart_repeats_the_run():
    """Rounds with restart: a system whose run ended starts again from its
    initial record; run on to the end without restart, every system's
    outputs equal the whole run's."""
    print("gpu: rounds with restart", flush=True)
    specs = _specs(count=16, seed=33)
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    b.run()
    ref = b.download()
    b.reset()
    for _ in range(4):  # 4 x 200 steps: every run (<= 615 steps) ends, the short ones restart
        b.round(200, restart=True)
    while b.round(200, restart=False)["system_steps"] > 0:
        pass
    for a, c in zip(b.download(), ref):
        _same(a, c)
    b.close()


def test_autotune_round_keeps_results():
    print("gpu: autotune over rounds", flush=True)
    specs = _specs(count=12, seed=34)
    states = W.initial_states(specs, processes=1)
    b = Batch(specs, states)
    b.run()
    ref = b.download()
    tune = b.autotune_round(100)
    assert 0 <= tune["best"] < len(tune["candidates"])
    assert all(c["ms"] > 0 for c in tune["candidates"])
    b.run()  # the tuner reset the batch: a whole run from the start
    for a, c in zip(b.download(), ref):
        _same(a, c)
    from nbodygradient_amd import _lib
    ctx = _lib.default_context()
    ctx.set_coop(1, 128)
    ctx.set_launch(0)
    ctx.set_systems_per_wave(1)
    b.close()
