"""The device build of the portable libm (nbg_libm.h, through
nbg_libm_eval_device) equals the host build (the oracle's) bit for bit over
the whole double range: the Cody-Waite and Payne-Hanek branches of sin/cos,
exp to its overflow and underflow ends, sinh/cosh on both sides of their
series switch, cbrt over normals and subnormals, and the special values.
The host build's accuracy against binary128 is pinned in
test_oracle_utils.py, so this carries it to the device."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from nbodygradient_amd import _lib

pytestmark = pytest.mark.gpu

FN = {"sin": 0, "cos": 1, "exp": 2, "sinh": 3, "cosh": 4, "cbrt": 5}
SPECIALS = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.0 ** -1022, np.finfo(float).max,
            -np.finfo(float).max, 1.0, -1.0, 0.5, 2.0 ** -27, 2.0 ** -28, 6381956970095103 * 2.0 ** 797]


def _device(name, x):
    ctx = _lib.default_context()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    ctx.check(_lib.lib().nbg_libm_eval_device(ctx.handle, FN[name], len(x), x.ctypes.data_as(_lib.D),
                                              out.ctypes.data_as(_lib.D)), "nbg_libm_eval_device")
    return out


def _inputs(name, rng):
    if name in ("sin", "cos"):
        parts = [rng.uniform(-20, 20, 30000), rng.uniform(-1.6e6, 1.6e6, 30000),
                 10 ** rng.uniform(-30, 308, 30000), rng.uniform(-2.0 ** -26, 2.0 ** -26, 2000)]
    elif name == "exp":
        parts = [rng.uniform(-750, 715, 60000), rng.uniform(-1, 1, 20000)]
    elif name in ("sinh", "cosh"):
        parts = [rng.uniform(-712, 712, 40000), rng.uniform(-1.2, 1.2, 40000)]
    else:
        parts = [10 ** rng.uniform(-320, 308, 60000), rng.uniform(0.1, 10, 20000)]
    x = np.concatenate(parts + [np.array(SPECIALS)])
    x[1:-len(SPECIALS):2] *= -1
    return x


@pytest.mark.parametrize("name", list(FN))
def test_device_libm_equals_host_build(name):
    print("gpu: device libm", name, flush=True)
    x = _inputs(name, np.random.default_rng(100 + FN[name]))
    dev = _device(name, x)
    host, _ = O.libm_eval(name, x)
    same = (dev.view(np.int64) == host.view(np.int64)) | (np.isnan(dev) & np.isnan(host))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, (name, x[bad[:5]], dev[bad[:5]], host[bad[:5]])


def test_device_libm_rejects_bad_arguments():
    ctx = _lib.default_context()
    x = np.zeros(4)
    assert _lib.lib().nbg_libm_eval_device(ctx.handle, 6, 4, x.ctypes.data_as(_lib.D), x.ctypes.data_as(_lib.D)) != 0
    assert _lib.lib().nbg_libm_eval_device(ctx.handle, 0, 0, None, None) == 0
