"""Shared set-ups restated from the reference's test files (test/*.jl).

Each helper names the reference test it follows.  Data come from the
reference's own fixtures (tests/golden/elements.txt = test/elements.txt,
tests/golden/coordinates.txt = test/coordinates.txt).
"""
import os

import numpy as np

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


def elements():
    return np.loadtxt(os.path.join(HERE, "golden", "elements.txt"), delimiter=",")


def coordinates():
    return np.loadtxt(os.path.join(HERE, "golden", "coordinates.txt"), delimiter=",")


def maxabs(a):
    return float(np.max(np.abs(a)))


def rel_maxabs(a, b):
    """isapprox(a, b; norm=maxabs) error: maxabs(a-b) / max(maxabs(a), maxabs(b))."""
    den = max(maxabs(a), maxabs(b))
    return maxabs(np.asarray(a) - np.asarray(b)) / den if den > 0 else maxabs(np.asarray(a) - np.asarray(b))


def tilt(x, v, nb):
    """perturb!/tilt in test_integrator.jl:18-26 and test_phisalpha.jl:35-40."""
    x = x.copy(order="F")
    v = v.copy(order="F")
    x[1, 0] = 5e-1 * np.sqrt(x[0, 0] ** 2 + x[2, 0] ** 2)
    x[1, 1] = -5e-1 * np.sqrt(x[0, 1] ** 2 + x[2, 1] ** 2)
    v[1, 0] = 5e-1 * np.sqrt(v[0, 0] ** 2 + v[2, 0] ** 2)
    v[1, 1] = -5e-1 * np.sqrt(v[0, 1] ** 2 + v[2, 1] ** 2)
    if nb >= 3:
        x[1, 2] = -5e-1 * np.sqrt(x[0, 1] ** 2 + x[2, 1] ** 2)
        v[1, 2] = -5e-1 * np.sqrt(v[0, 1] ** 2 + v[2, 1] ** 2)
    return x, v


def integrator_case():
    """test_integrator.jl:5-29: n=3, inner masses x100, Omega = 0, tilted."""
    el = elements()[:3].copy()
    el[1, 0] *= 100.0
    el[2, 0] *= 100.0
    el[:, -1] = 0.0
    t0 = 7257.93115525
    x, v, jac_init = O.init_nbody(t0, el)
    x, v = tilt(x, v, 3)
    return el, t0, x, v, jac_init


def transit_case():
    """test_transit_timing.jl:4-16: n=3, t0 before first transit, masses x10."""
    el = elements()[:3].copy()
    el[1:, 2] -= 7300.0
    el[:, -1] = 0.0
    el[1, 0] *= 10.0
    el[2, 0] *= 10.0
    t0 = 7257.93115525 - 7300.0 - 0.5
    return el, t0
