"""Initial conditions on the GPU (nbg_batch_create_elements): x, v and
jac_init from orbital elements, checked against the oracle's init_nbody
(itself pinned by finite differences and the TRAPPIST-1 fixtures).  The
device takes atan2 from its own math library and sin/cos from nbg_libm.h,
the oracle from glibc, so agreement is to rounding: x and v within 1e-13 of
each system's scale, jac_init within 1e-11 of each column's scale.  End to
end, batches started from device and host initial conditions give the same
transit counts, transit times within 1e-12 relative and dtdelements within
1e-10 of each block's scale."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
import nbodygradient_amd as nb
from nbodygradient_amd import _lib
from nbodygradient_amd import workloads as W
from nbodygradient_amd.batch import Batch

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _specs(seed, steps):
    rng = np.random.default_rng(seed)
    out = []
    for n in range(2, 13):
        s = W.c2_system(rng, n)
        s.tmax = min(s.tmax, steps * s.h)
        out.append(s)
    for s in W.c3(count=6, seed=seed)[:6]:
        s.tmax = steps * s.h
        out.append(s)
    return out


def test_device_ic_matches_oracle_init_nbody():
    say("gpu: device IC vs oracle init_nbody")
    specs = _specs(81, 50)
    b = Batch(specs, device_ic=True)
    got = b.initial()
    b.run()  # other kernels in between: nothing may depend on what LDS held before
    b.close()
    b = Batch(specs, device_ic=True)
    again = b.initial()
    b.close()
    for g, a in zip(got, again):
        assert all(np.array_equal(p, q) for p, q in zip(g, a))
    worst_x = worst_j = 0.0
    for s, (x, v, J) in zip(specs, got):
        ox, ov, oJ = O.init_nbody(s.t0, s.elements[:, :7])
        sx, sv = np.abs(ox).max(), np.abs(ov).max()
        ex = max(np.abs(x - ox).max() / sx, np.abs(v - ov).max() / sv)
        col = np.maximum(np.abs(oJ).max(axis=0), 1e-300)
        ej = (np.abs(J - oJ).max(axis=0) / col).max()
        worst_x, worst_j = max(worst_x, ex), max(worst_j, ej)
        assert ex < 1e-13, (s.n, ex)
        assert ej < 1e-11, (s.n, ej)
        # exact structure: unit mass rows
        for i in range(s.n):
            row = np.zeros(7 * s.n)
            row[7 * i + 6] = 1.0
            assert np.array_equal(J[7 * i + 6, :], row)
    say("  worst relative difference: x/v %.2e, jac_init %.2e" % (worst_x, worst_j))


def test_device_ic_batch_transits_match_host_ic_batch():
    say("gpu: device-IC batch vs host-IC batch")
    specs = _specs(82, 1500)
    bd = Batch(specs, device_ic=True)
    bd.run()
    rd = bd.download(derivatives=True)
    bd.close()
    states = W.initial_states(specs, processes=1)
    bh = Batch(specs, states)
    bh.run()
    rh = bh.download(derivatives=True)
    bh.close()
    worst_t = worst_d = 0.0
    for s, a, h in zip(specs, rd, rh):
        assert np.array_equal(a["count"], h["count"]), s.n
        for i in range(s.n):
            k = int(min(h["count"][i], s.ntt()))
            if k == 0:
                continue
            tt_a, tt_h = a["tt"][i, :k], h["tt"][i, :k]
            et = (np.abs(tt_a - tt_h) / np.maximum(np.abs(tt_h), 1.0)).max()
            worst_t = max(worst_t, et)
            assert et < 1e-12, (s.n, i, et)
            da, dh = a["dtdelements"][i, :k], h["dtdelements"][i, :k]
            ed = np.abs(da - dh).max() / max(np.abs(dh).max(), 1e-300)
            worst_d = max(worst_d, ed)
            assert ed < 1e-10, (s.n, i, ed)
    say("  worst relative difference: tt %.2e, dtdelements %.2e" % (worst_t, worst_d))


def test_device_ic_rejects_a_hierarchy_that_does_not_fit():
    say("gpu: device IC hierarchy check")
    ctx = nb.default_context()
    n = 3
    el = np.asfortranarray(W.c2_system(np.random.default_rng(83), n).elements[:, :7])
    # the second orbit is a binary without body 1 (a new binary): the
    # reference would read elements row 4 of 3 (init_nbody.jl:58-75)
    H = np.asfortranarray(np.array([[-1.0, 1.0, 0.0], [0.0, -1.0, 1.0], [-1.0, -1.0, -1.0]]))
    h = C.c_void_p()
    args = [ctx.handle, 1, np.array([n], dtype=np.int32).ctypes.data_as(_lib.I32),
            np.array([0.1]).ctypes.data_as(_lib.D), np.array([10.0]).ctypes.data_as(_lib.D),
            np.array([0.0]).ctypes.data_as(_lib.D), None, np.array([5], dtype=np.int64).ctypes.data_as(_lib.I64),
            1, 1, el.ctypes.data_as(_lib.D), H.ctypes.data_as(_lib.D), None, C.byref(h)]
    rc = _lib.lib().nbg_batch_create_elements(*args)
    assert rc != 0 and b"hierarchy" in _lib.lib().nbg_last_error(ctx.handle)
