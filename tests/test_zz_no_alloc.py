"""nbg_integrate_transits_noalloc (the entry point that allocates no device
memory: its record lives in the library's fixed scratch array) against
nbg_integrate_transits, through the C ABI, bit for bit: two calls in a row,
the second with a smaller system, so that the scratch array's leftovers from
the first cannot leak into the second; with and without gradients; and a
system too large for the array is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from nbodygradient_amd import _lib, workloads as W
from nbodygradient_amd._lib import dptr, i64ptr, u8ptr
from nbodygradient_amd.ics import ElementsIC, init_nbody

pytestmark = pytest.mark.gpu

D = C.POINTER(C.c_double)


def _bind():
    L = _lib.lib()
    L.nbg_integrate_transits_noalloc.argtypes = L.nbg_integrate_transits.argtypes
    return L


def _record(spec, grad, ntbv=1, ti=0):
    n = spec.n
    x, v, jinit = init_nbody(ElementsIC(spec.t0, n, spec.elements))
    ntt = spec.ntt()
    shp_tt = (n, ntt) if ntbv == 1 else (3, n, ntt)
    shp_dq = (n, ntt, 7, n) if ntbv == 1 else (3, n, ntt, 7, n)
    rng = np.random.default_rng(n)
    return {"t": C.c_double(spec.t0), "x": x.copy(order="F"), "v": v.copy(order="F"),
            "xe": np.zeros((3, n), order="F"), "ve": np.zeros((3, n), order="F"),
            "m": np.ascontiguousarray(spec.elements[:, 0]), "pair": np.zeros((n, n), dtype=np.uint8, order="F"),
            "J": np.asfortranarray(np.eye(7 * n)), "JE": np.zeros((7 * n, 7 * n), order="F"),
            "dqdt": rng.standard_normal(7 * n), "jinit": jinit.copy(order="F"), "ntt": ntt,
            # outputs start from the same arbitrary contents (entries beyond
            # count must stay as given)
            "tt": np.asfortranarray(rng.standard_normal(shp_tt)),
            "dq": np.asfortranarray(rng.standard_normal(shp_dq)),
            "de": np.asfortranarray(rng.standard_normal(shp_dq)),
            "count": np.zeros(n, dtype=np.int64), "n": n, "h": spec.h, "tmax": spec.tmax, "grad": grad,
            "ntbv": ntbv, "ti": ti}


def _call(fn, r):
    ctx = _lib.default_context()
    rc = fn(ctx.handle, r["n"], C.c_double(r["h"]), C.c_double(r["tmax"]), r["ti"], int(r["grad"]), r["ntbv"],
            C.byref(r["t"]), dptr(r["x"]), dptr(r["v"]), dptr(r["xe"]), dptr(r["ve"]), dptr(r["m"]), u8ptr(r["pair"]),
            dptr(r["J"]), dptr(r["JE"]), dptr(r["dqdt"]), dptr(r["jinit"]), C.c_int64(r["ntt"]), dptr(r["tt"]),
            dptr(r["dq"]), dptr(r["de"]), i64ptr(r["count"]))
    return rc, (_lib.lib().nbg_last_error(ctx.handle).decode() if rc else "")


def _same(a, b):
    assert a["t"].value == b["t"].value
    for k in ("x", "v", "xe", "ve", "J", "JE", "dqdt", "tt", "dq", "de", "count"):
        assert np.array_equal(a[k], b[k]), k


def _specs():
    rng = np.random.default_rng(41)
    big = W.c2_system(rng, 8)
    big.tmax = 700 * big.h
    small = W.c2_system(rng, 3)
    small.tmax = 400 * small.h
    return big, small


@pytest.mark.parametrize("grad", [True, False])
def test_noalloc_equals_allocating_entry_point(grad):
    print("gpu: no-alloc entry point grad=%s" % grad, flush=True)
    L = _bind()
    for spec, ntbv in zip(_specs(), (1, 3)):  # the second call is smaller
        a = _record(spec, grad, ntbv)
        b = _record(spec, grad, ntbv)
        rc, msg = _call(L.nbg_integrate_transits, a)
        assert rc == 0, msg
        rc, msg = _call(L.nbg_integrate_transits_noalloc, b)
        assert rc == 0, msg
        assert int(b["count"].sum()) > 0
        _same(a, b)


def test_noalloc_rejects_a_system_beyond_its_scratch():
    L = _bind()
    rng = np.random.default_rng(42)
    spec = W.c2_system(rng, 10)
    spec.tmax = 100 * spec.h
    r = _record(spec, True)
    # dtdq0 and dtdelements alone: 2 x 8 x n x ntt x 7n bytes > 32 MiB
    r["ntt"] = 3000
    r["tt"] = np.zeros((10, 3000), order="F")
    r["dq"] = np.zeros((10, 3000, 7, 10), order="F")
    r["de"] = np.zeros((10, 3000, 7, 10), order="F")
    x0 = r["x"].copy()
    rc, msg = _call(L.nbg_integrate_transits_noalloc, r)
    assert rc == -1 and "scratch" in msg
    assert np.array_equal(r["x"], x0)
