"""The device pair formulas (nbodygradient.jl_amd/csrc/nbg_pair.h) against
the oracle, bit for bit, on the CPU.

libnbg_hip.so exports nbg_selftest_pair_host, which runs the very same
header compiled for the host; no GPU is involved.  Inputs sweep elliptic and
hyperbolic relative orbits, both drift orders, small and large steps (series
and closed-form branches of G3/H*), and near-degenerate zeta.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        r = 10 ** rng.uniform(-2, 1)
        x0 = rng.normal(size=3)
        x0 *= r / np.linalg.norm(x0)
        k = 10 ** rng.uniform(-9, -3)          # G (m_i + m_j)
        vc = np.sqrt(k / r)
        v0 = rng.normal(size=3)
        v0 *= vc * rng.uniform(0.2, 2.0) / np.linalg.norm(v0)  # beta of either sign
        h = 10 ** rng.uniform(-3, 1.5) * rng.choice([-1.0, 1.0])
        yield x0, v0, k, h


@pytest.mark.parametrize("drift_first", [True, False])
def test_pair_formulas_bitwise(drift_first):
    L = nb.lib()
    D = C.POINTER(C.c_double)
    nb_out = np.zeros(61)
    orc_out = np.zeros(61)
    nbad = 0
    for x0, v0, k, h in _cases(42 + int(drift_first), 3000):
        x0 = np.ascontiguousarray(x0)
        v0 = np.ascontiguousarray(v0)
        rc = L.nbg_selftest_pair_host(x0.ctypes.data_as(D), v0.ctypes.data_as(D), k, h, int(drift_first),
                                      nb_out.ctypes.data_as(D))
        assert rc == 0
        O.lib().orc_pair_kepler(x0.ctypes.data_as(D), v0.ctypes.data_as(D), C.c_double(k), C.c_double(h),
                                int(drift_first), orc_out.ctypes.data_as(D))
        same = np.array_equal(nb_out, orc_out) or np.array_equal(
            np.nan_to_num(nb_out, nan=7.0), np.nan_to_num(orc_out, nan=7.0))
        nbad += not same
    assert nbad == 0
