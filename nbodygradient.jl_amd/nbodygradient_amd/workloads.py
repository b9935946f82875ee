"""Seeded synthetic systems for the benchmark configurations.

Every system is a star (m = 1) plus planets on nearly edge-on orbits, given as
the reference's element rows (m, P, t0, e cos w, e sin w, I, Omega) with t0 the
time of transit; drawing t0 uniformly over one period is the same as drawing
the mean anomaly at epoch uniformly.

  c0: 1 system, 2 planets, P = 10.0 and 21.3 d, h = P_in/40, tmax = 1000 d
  c1: 1 system, 7 planets, P from 1.5 d with ratios U[1.49, 1.55], h = 0.06, tmax = 4000 d
  c2: 4096 systems, n ~ U{3..9} bodies, P log-uniform in [2, 200] d (sorted,
      ratios >= 1.3), h = P_in/30, tmax ~ U[500, 4000] d
  c3: 16384 Gaussian perturbations of one base system with P = 5, 8.1, 13.2,
      21.6, 35.0 d (star + 5 planets), h = 0.1, tmax = 3000 d
  c4: request stream, Poisson arrivals, n ~ U{2..10}, tmax log-uniform in
      [100, 8000] d, elements as c2
"""
import math
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .ics import ElementsIC, init_nbody

TWO_PI = 2.0 * math.pi


class SystemSpec:
    __slots__ = ("elements", "t0", "h", "tmax")

    def __init__(self, elements, t0, h, tmax):
        self.elements = np.asarray(elements, dtype=np.float64)
        self.t0 = float(t0)
        self.h = float(h)
        self.tmax = float(tmax)

    @property
    def n(self):
        return self.elements.shape[0]

    @property
    def nsteps(self):
        return abs(int(np.rint(self.tmax / self.h)))

    def ntt(self):
        P = self.elements[1:, 1]
        return int(np.max(np.ceil(np.abs(self.tmax / P)).astype(np.int64) + 3))


def _planet(rng, m_lo, m_hi, P, e_hi, inc_dev, Omega=None):
    m = math.exp(rng.uniform(math.log(m_lo), math.log(m_hi)))
    e = rng.uniform(0.0, e_hi)
    w = rng.uniform(0.0, TWO_PI)
    t0 = rng.uniform(0.0, P)
    inc = math.pi / 2 + rng.uniform(-inc_dev, inc_dev)
    Om = rng.uniform(0.0, TWO_PI) if Omega is None else Omega
    return [m, P, t0, e * math.cos(w), e * math.sin(w), inc, Om]


def c0(seed=20261202):
    rng = np.random.default_rng(seed)
    rows = [[1.0, 0, 0, 0, 0, 0, 0]]
    for P in (10.0, 21.3):
        rows.append(_planet(rng, 1e-6, 3e-5, P, 0.05, 0.01))
    return SystemSpec(rows, 0.0, 10.0 / 40.0, 1000.0)


def c1(seed=20261203):
    rng = np.random.default_rng(seed)
    rows = [[1.0, 0, 0, 0, 0, 0, 0]]
    P = 1.5
    for _ in range(7):
        rows.append(_planet(rng, 1e-6, 1e-5, P, 0.02, 0.005, Omega=rng.uniform(-0.005, 0.005)))
        P *= rng.uniform(1.49, 1.55)
    return SystemSpec(rows, 0.0, 0.06, 4000.0)


def _periods(rng, k, lo=2.0, hi=200.0, min_ratio=1.3):
    while True:
        P = np.sort(np.exp(rng.uniform(math.log(lo), math.log(hi), k)))
        if k == 1 or np.all(P[1:] / P[:-1] >= min_ratio):
            return P


def c2_system(rng, n):
    P = _periods(rng, n - 1)
    rows = [[1.0, 0, 0, 0, 0, 0, 0]] + [_planet(rng, 1e-6, 3e-5, p, 0.05, 0.01) for p in P]
    tmax = rng.uniform(500.0, 4000.0)
    return SystemSpec(rows, 0.0, P[0] / 30.0, tmax)


def c2(count=4096, seed=20261204):
    rng = np.random.default_rng(seed)
    return [c2_system(rng, int(rng.integers(3, 10))) for _ in range(count)]


def c3(count=16384, seed=20261205):
    rng = np.random.default_rng(seed)
    base = [[1.0, 0, 0, 0, 0, 0, 0]] + [_planet(rng, 1e-6, 3e-5, P, 0.05, 0.01)
                                         for P in (5.0, 8.1, 13.2, 21.6, 35.0)]
    base = np.array(base)
    out = []
    for _ in range(count):
        el = base.copy()
        el[1:, 0] *= 1.0 + 1e-4 * rng.standard_normal(el.shape[0] - 1)
        el[1:, 1] *= 1.0 + 1e-5 * rng.standard_normal(el.shape[0] - 1)
        el[1:, 3] *= 1.0 + 1e-3 * rng.standard_normal(el.shape[0] - 1)
        el[1:, 4] *= 1.0 + 1e-3 * rng.standard_normal(el.shape[0] - 1)
        out.append(SystemSpec(el, 0.0, 0.1, 3000.0))
    return out


def c4(count=2000, rate=2000.0, seed=20261206):
    """(arrival_time_s, SystemSpec) pairs with Poisson arrivals."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(2, 11))
        spec = c2_system(rng, n)
        spec.tmax = math.exp(rng.uniform(math.log(100.0), math.log(8000.0)))
        out.append((t, spec))
    return out


def _ic_one(spec):
    ic = ElementsIC(spec.t0, spec.n, spec.elements)
    return init_nbody(ic)


def initial_states(specs, processes=None):
    """init_nbody for every system (x, v, jac_init), in parallel processes."""
    if processes is None:
        processes = min(16, os.cpu_count() or 1)
    if processes <= 1 or len(specs) < 64:
        return [_ic_one(s) for s in specs]
    with ProcessPoolExecutor(processes) as ex:
        return list(ex.map(_ic_one, specs, chunksize=max(1, len(specs) // (4 * processes))))
