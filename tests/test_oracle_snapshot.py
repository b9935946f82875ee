"""The oracle's TransitSnapshot run (Transits.jl:186-211, snapshot.jl:7-54),
pinned by the reference's own test (test_transit_snapshot.jl: TRAPPIST-1
with three bodies, masses x10): at the transit times TransitParameters finds,
the snapshot's v_sky and b^2 equal the transit search's within 1e-10.  The
snapshot's derivatives there agree with the transit search's dtbvdq0 (both
apply the same chain rule to J, by different integration paths: full steps
plus a partial one, against the prior state plus the Newton step)."""
import os

import numpy as np

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _case():
    el = np.loadtxt(os.path.join(ROOT, "tests", "golden", "elements.txt"), delimiter=",")[:3].copy()
    el[1:, 2] -= 7300.0
    el[:, -1] = 0.0
    el[1, 0] *= 10.0
    el[2, 0] *= 10.0
    t0 = 7257.93115525 - 7300.0 - 0.5
    return el, t0, 0.04, 10.0


def _state(el, t0):
    x, v, jinit = O.init_nbody(t0, el)
    return O.OState(x, v, el[:, 0], t=t0, jac_init=jinit)


def test_snapshot_at_transit_times_reproduces_transit_parameters():
    el, t0, h, tmax = _case()
    P = el[:, 1]
    ntt = int(np.max(np.ceil(np.abs(tmax / P[1:])))) + 3
    s = _state(el, t0)
    ttbv, dq, _, count = O.integrate_transits(s, h, tmax, ntt, ti=0, grad=True, ntbv=3)
    k = int(count[1])
    assert k >= 2
    times = ttbv[0, 1, :k]
    s2 = _state(el, t0)
    vsky, bsky2, dbv, dbvel = O.transit_snapshot(s2, h, times, grad=True)
    # test_transit_snapshot.jl: tol = 1e-10 on v_sky and b^2 of body 2
    assert np.all(np.abs(ttbv[1, 1, :k] - vsky[1, :]) < 1e-10)
    assert np.all(np.abs(ttbv[2, 1, :k] - bsky2[1, :]) < 1e-10)
    # derivatives: dtbvdq0[1:3] (v_sky, b^2) against the snapshot's dbvdq0.
    # v_sky: every entry above a tenth of its block's largest.  b^2: this
    # system is edge-on (I = pi/2, Omega = 0), so b^2 ~ 1e-36 and its
    # derivative is 2 (dJ_x dx + dJ_y dy) with the sky separations dx, dy
    # near zero; dx is as small as each path's last digits make it, so only
    # the entries through dy (the y rows, k = 2 and 5) are comparable.
    for j in range(k):
        a = dbv[0, 1, j]
        b = dq[1, 1, j]
        big = np.abs(b) > 0.1 * np.abs(b).max()
        assert big.sum() >= 3
        assert (np.abs(a - b)[big] / np.abs(b)[big]).max() < 1e-5, j  # paths differ in their last steps
        a = dbv[1, 1, j, [1, 4], :]
        b = dq[2, 1, j, [1, 4], :]
        big = np.abs(b) > 1e-3 * np.abs(b).max()
        assert (np.abs(a - b)[big] / np.abs(b)[big]).max() < 1e-5, j
    # body 1 is the reference body: its entries stay zero
    assert not vsky[0].any() and not dbv[:, 0].any()
    # dbvdelements = dbvdq0 . jac_init, element by element
    J = s2.jac_init
    n = 3
    Jr = J.reshape((7, n, 7, n), order="F")  # Jr[q, p, l, k] = jac_init[7p + q, 7k + l]
    ref = np.einsum("aijqp,qplk->aijlk", dbv, Jr)
    assert np.allclose(dbvel, ref, rtol=1e-12, atol=1e-300)


def test_snapshot_without_grad_gives_the_same_positions():
    el, t0, h, _ = _case()
    times = np.array([t0 + 0.7, t0 + 3.3, t0 + 3.3, t0 + 9.1])
    g = _state(el, t0)
    ng = _state(el, t0)
    vg, bg, _, _ = O.transit_snapshot(g, h, times, grad=True)
    vn, bn, dq, de = O.transit_snapshot(ng, h, times, grad=False)
    assert np.array_equal(vg, vn) and np.array_equal(bg, bn)
    assert not dq.any() and not de.any()
    assert g.t == times[-1] and np.array_equal(g.x, ng.x)
    # a repeated time integrates nothing more: the same snapshot twice
    assert np.array_equal(vg[:, 1], vg[:, 2])
