"""Writes tests/golden/c1_oracle.npz: the oracle's full run of the 8-body
near-resonant system (star + 7 planets, P from 1.5 d with ratios
U[1.49, 1.55], h = 0.06 d, tmax = 4000 d: 66,667 steps), transit times and
counts in full, and the dtdelements rows of every 25th transit of every
body (the whole array is 9.6 MB).  Run from the repository root after
building the oracle (make -C oracle):

    python tests/golden/make_c1_golden.py

The run takes about a minute on one core; the GPU test compares the GPU's
full run with this file."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "nbodygradient.jl_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from nbodygradient_amd import workloads as W  # noqa: E402

EVERY = 25


def main():
    spec = W.c1()
    x, v, jinit = W.initial_states([spec], processes=1)[0]
    s = O.OState(x, v, spec.elements[:, 0], t=spec.t0, jac_init=jinit)
    tt, dq, de, count = O.integrate_transits(s, spec.h, spec.tmax, spec.ntt(), grad=True)
    sel = np.arange(0, spec.ntt(), EVERY)
    np.savez_compressed(os.path.join(HERE, "c1_oracle.npz"), tt=tt, count=count, every=EVERY,
                        dtdelements_sel=de[:, sel], x=s.x, v=s.v, nsteps=spec.nsteps)
    print("c1: %d steps, counts %s" % (spec.nsteps, count.tolist()))


if __name__ == "__main__":
    main()
