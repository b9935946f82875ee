"""ctypes access to the CPU oracle (oracle/build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.  Arrays follow the reference's Julia layout
(column-major, "F" order in numpy) with 0-based indices.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

_lib = None

GNEWT = 39.4845 / (365.242 * 365.242)

D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_longlong)
U8 = C.POINTER(C.c_uint8)
I32 = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_batch_transits.restype = C.c_double
    return _lib


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), a.flags
    return a.ctypes.data_as(D)


def ip(a):
    assert a.dtype == np.int64
    return a.ctypes.data_as(I64)


def up(a):
    if a is None:
        return None
    assert a.dtype == np.uint8
    return a.ctypes.data_as(U8)


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


# ------------------------------------------------------------------ libm
LIBM = {"sin": 0, "cos": 1, "exp": 2, "sinh": 3, "cosh": 4, "cbrt": 5}


def libm_eval(name, x):
    x = f64(x).ravel()
    out = np.zeros_like(x)
    ulp = np.zeros_like(x)
    lib().orc_libm_eval(LIBM[name], len(x), dp(x), dp(out), dp(ulp))
    return out, ulp


UTILS = {"cubic1": 0, "G3": 1, "H1": 2, "H2": 3, "H3": 4, "H5": 5, "H6": 6,
         "G3_series": 7, "H1_series": 8, "H2_series": 9, "H3_series": 10, "H5_series": 11, "H6_series": 12}


def utils_eval(name, a, b, c=None, quad=False):
    a = f64(a).ravel()
    b = f64(b).ravel()
    c = f64(c).ravel() if c is not None else np.zeros_like(a)
    out = np.zeros_like(a)
    fn = lib().orc_utils_eval_q if quad else lib().orc_utils_eval
    fn(UTILS[name], len(a), dp(a), dp(b), dp(c), dp(out))
    return out


def comp_sum(sv, se, add):
    sv = f64(sv).copy()
    se = f64(se).copy()
    add = f64(add)
    lib().orc_comp_sum(len(sv), dp(sv), dp(se), dp(add))
    return sv, se


# ------------------------------------------------------------------ ICs
def hierarchy_nested(n):
    H = np.zeros((n, n), order="F")
    lib().orc_hierarchy_nested(n, dp(H))
    return H


def init_nbody(t0, elements, H=None):
    el = f64(elements)
    n = el.shape[0]
    if H is None:
        H = hierarchy_nested(n)
    H = f64(H)
    x = np.zeros((3, n), order="F")
    v = np.zeros((3, n), order="F")
    jac = np.zeros((7 * n, 7 * n), order="F")
    lib().orc_init_nbody(n, C.c_double(t0), dp(el), dp(H), dp(x), dp(v), dp(jac))
    return x, v, jac


def kepler_init(t, mass, el):
    el = f64(el)
    x = np.zeros(3)
    v = np.zeros(3)
    jac = np.zeros((7, 7), order="F")
    lib().orc_kepler_init(C.c_double(t), C.c_double(mass), dp(el), dp(x), dp(v), dp(jac))
    return x, v, jac


def fd_kepler_init(t, mass, el, dq=1e-15):
    el = f64(el)
    jac = np.zeros((7, 7), order="F")
    lib().orc_fd_kepler_init(C.c_double(t), C.c_double(mass), dp(el), C.c_double(dq), dp(jac))
    return jac


def fd_init_nbody(t0, elements, H=None, dq=1e-15):
    el = f64(elements)
    n = el.shape[0]
    if H is None:
        H = hierarchy_nested(n)
    H = f64(H)
    jac = np.zeros((7 * n, 7 * n), order="F")
    lib().orc_fd_init_nbody(n, C.c_double(t0), dp(el), dp(H), C.c_double(dq), dp(jac))
    return jac


# ------------------------------------------------------------------ operators
def kepler_driftij(x, v, m, i, j, h, drift_first):
    x = f64(x).copy(order="F")
    v = f64(v).copy(order="F")
    m = f64(m)
    n = x.shape[1]
    jac = np.zeros((14, 14), order="F")
    dq = np.zeros(14)
    lib().orc_kepler_driftij(n, dp(x), dp(v), dp(m), i, j, C.c_double(h), int(drift_first), dp(jac), dp(dq))
    return x, v, jac, dq


def fd_kepler_driftij(x, v, m, i, j, h, drift_first, dlnq=1e-15):
    x = f64(x); v = f64(v); m = f64(m)
    n = x.shape[1]
    jac = np.zeros((14, 14), order="F")
    dq = np.zeros(14)
    lib().orc_fd_kepler_driftij(n, dp(x), dp(v), dp(m), i, j, C.c_double(h), int(drift_first),
                                C.c_double(dlnq), dp(jac), dp(dq))
    return jac, dq


OPS = {"kickfast": 0, "phic": 1, "phisalpha": 2}


def operator(name, x, v, m, pair, h):
    x = f64(x).copy(order="F"); v = f64(v).copy(order="F"); m = f64(m)
    n = x.shape[1]
    pair = np.asfortranarray(np.asarray(pair, dtype=np.uint8))
    jac = np.zeros((7 * n, 7 * n), order="F")
    dq = np.zeros(7 * n)
    lib().orc_operator(OPS[name], n, dp(x), dp(v), dp(m), up(pair), C.c_double(h), dp(jac), dp(dq))
    return x, v, jac, dq


def fd_operator(name, x, v, m, pair, h, dlnq=1e-15):
    x = f64(x); v = f64(v); m = f64(m)
    n = x.shape[1]
    pair = np.asfortranarray(np.asarray(pair, dtype=np.uint8))
    jac = np.zeros((7 * n, 7 * n), order="F")
    dq = np.zeros(7 * n)
    lib().orc_fd_operator(OPS[name], n, dp(x), dp(v), dp(m), up(pair), C.c_double(h), C.c_double(dlnq), dp(jac), dp(dq))
    return jac, dq


# ------------------------------------------------------------------ state
class OState:
    """Plain container with the reference State's integration fields."""

    def __init__(self, x, v, m, t=0.0, jac=None, jac_init=None, pair=None):
        x = f64(x)
        n = x.shape[1]
        self.n = n
        self.x = x.copy(order="F")
        self.v = f64(v).copy(order="F")
        self.m = f64(m).copy()
        self.t = float(t)
        self.xerror = np.zeros((3, n), order="F")
        self.verror = np.zeros((3, n), order="F")
        self.jac_step = np.asfortranarray(np.eye(7 * n)) if jac is None else f64(jac).copy(order="F")
        self.jac_error = np.zeros((7 * n, 7 * n), order="F")
        self.dqdt = np.zeros(7 * n)
        self.jac_init = np.asfortranarray(np.eye(7 * n)) if jac_init is None else f64(jac_init).copy(order="F")
        self.pair = np.zeros((n, n), dtype=np.uint8, order="F") if pair is None else np.asfortranarray(pair.astype(np.uint8))

    def copy(self):
        o = OState(self.x, self.v, self.m, self.t, self.jac_step, self.jac_init, self.pair)
        o.xerror = self.xerror.copy(order="F")
        o.verror = self.verror.copy(order="F")
        o.jac_error = self.jac_error.copy(order="F")
        o.dqdt = self.dqdt.copy()
        return o


def step(s, h, grad=True):
    lib().orc_ahl21_step(s.n, C.c_double(h), int(grad), dp(s.x), dp(s.v), dp(s.xerror), dp(s.verror), dp(s.m),
                         up(s.pair), dp(s.jac_step), dp(s.jac_error), dp(s.dqdt))
    return s


def integrate_steps(s, h, N, grad=True):
    t = C.c_double(s.t)
    lib().orc_integrate(0, s.n, C.c_double(h), C.c_longlong(N), C.c_double(0.0), int(grad), C.byref(t),
                        dp(s.x), dp(s.v), dp(s.xerror), dp(s.verror), dp(s.m), up(s.pair),
                        dp(s.jac_step), dp(s.jac_error), dp(s.dqdt))
    s.t = t.value
    return s


def cartesian_output(s, h, tmax, nstep, grad=True, jacobian=True):
    """(intr)(s, o::CartesianOutput) (Outputs.jl:26-49) on the oracle: the
    state before each step as arrays over the step index (last axis)."""
    n = s.n
    rec = dict(t=np.zeros(nstep), x=np.zeros((3, n, nstep), order="F"), v=np.zeros((3, n, nstep), order="F"),
               xerror=np.zeros((3, n, nstep), order="F"), verror=np.zeros((3, n, nstep), order="F"))
    if jacobian:
        rec["jac_step"] = np.zeros((7 * n, 7 * n, nstep), order="F")
        rec["jac_error"] = np.zeros((7 * n, 7 * n, nstep), order="F")
    t = C.c_double(s.t)
    lib().orc_cartesian(s.n, C.c_double(h), C.c_double(tmax), C.c_longlong(nstep), int(grad), C.byref(t),
                        dp(s.x), dp(s.v), dp(s.xerror), dp(s.verror), dp(s.m), up(s.pair), dp(s.jac_step),
                        dp(s.jac_error), dp(s.dqdt), dp(rec["t"]), dp(rec["x"]), dp(rec["v"]), dp(rec["xerror"]),
                        dp(rec["verror"]), dp(rec.get("jac_step")), dp(rec.get("jac_error")))
    s.t = t.value
    return rec


def transit_snapshot(s, h, times, grad=True):
    """(intr)(s, ts::TransitSnapshot) (Transits.jl:186-211) on the oracle:
    vsky, bsky2 (n x nt), dbvdq0 and dbvdelements (2 x n x nt x 7 x n)."""
    n = s.n
    times = np.ascontiguousarray(times, dtype=np.float64)
    nt = len(times)
    vsky = np.zeros((n, nt), order="F")
    bsky2 = np.zeros((n, nt), order="F")
    dq = np.zeros((2, n, nt, 7, n), order="F")
    de = np.zeros((2, n, nt, 7, n), order="F")
    t = C.c_double(s.t)
    lib().orc_snapshot(n, C.c_double(h), nt, times.ctypes.data_as(D), int(grad), C.byref(t), dp(s.x), dp(s.v),
                       dp(s.xerror), dp(s.verror), dp(s.m), up(s.pair), dp(s.jac_step), dp(s.jac_error), dp(s.dqdt),
                       dp(s.jac_init), dp(vsky), dp(bsky2), dp(dq), dp(de))
    s.t = t.value
    return vsky, bsky2, dq, de


def integrate_to(s, h, time, grad=True):
    t = C.c_double(s.t)
    lib().orc_integrate(1, s.n, C.c_double(h), C.c_longlong(0), C.c_double(time), int(grad), C.byref(t),
                        dp(s.x), dp(s.v), dp(s.xerror), dp(s.verror), dp(s.m), up(s.pair),
                        dp(s.jac_step), dp(s.jac_error), dp(s.dqdt))
    s.t = t.value
    return s


def fd_steps(x, v, m, h, N, pair=None, dlnq=1e-20):
    x = f64(x); v = f64(v); m = f64(m)
    n = x.shape[1]
    jac = np.zeros((7 * n, 7 * n), order="F")
    pr = None if pair is None else np.asfortranarray(pair.astype(np.uint8))
    lib().orc_integrate_steps_q_fd(n, C.c_double(h), C.c_longlong(N), dp(x), dp(v), dp(m), up(pr),
                                   C.c_double(dlnq), dp(jac))
    return jac


def fd_step_dqdt(x, v, m, h, pair=None, dlnq=1e-20):
    x = f64(x); v = f64(v); m = f64(m)
    n = x.shape[1]
    dq = np.zeros(7 * n)
    pr = None if pair is None else np.asfortranarray(pair.astype(np.uint8))
    lib().orc_step_dqdt_q_fd(n, C.c_double(h), dp(x), dp(v), dp(m), up(pr), C.c_double(dlnq), dp(dq))
    return dq


def ntt_rule(tmax, elements):
    P = np.asarray(elements, dtype=np.float64)[:, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        q = tmax / P
    q = q[np.isfinite(q)]
    return int(np.max(np.ceil(np.abs(q)).astype(np.int64) + 3))


def integrate_transits(s, h, tmax, ntt, ti=0, grad=True, ntbv=1):
    n = s.n
    shape_tt = (n, ntt) if ntbv == 1 else (3, n, ntt)
    shape_dq = (n, ntt, 7, n) if ntbv == 1 else (3, n, ntt, 7, n)
    tt = np.zeros(shape_tt, order="F")
    dtdq0 = np.zeros(shape_dq, order="F")
    dtdel = np.zeros(shape_dq, order="F")
    count = np.zeros(n, dtype=np.int64)
    t = C.c_double(s.t)
    lib().orc_integrate_transits(n, C.c_double(h), C.c_double(tmax), ti, int(grad), ntbv, C.byref(t),
                                 dp(s.x), dp(s.v), dp(s.xerror), dp(s.verror), dp(s.m), up(s.pair),
                                 dp(s.jac_step), dp(s.jac_error), dp(s.dqdt), dp(s.jac_init), C.c_longlong(ntt),
                                 dp(tt), dp(dtdq0), dp(dtdel), ip(count))
    s.t = t.value
    return tt, dtdq0, dtdel, count


def fd_dtdelements(t0, elements, h, tmax, ti, ntt, H=None, dq0=1e-10):
    el = f64(elements)
    n = el.shape[0]
    if H is None:
        H = hierarchy_nested(n)
    H = f64(H)
    out = np.zeros((n, ntt, 7, n), order="F")
    lib().orc_fd_dtdelements(n, C.c_double(t0), dp(el), dp(H), C.c_double(h), C.c_double(tmax), ti,
                             C.c_double(dq0), C.c_longlong(ntt), dp(out))
    return out


def fd_dtbvdelements(t0, elements, h, tmax, ti, ntt, ntbv=3, H=None, dq0=1e-10):
    """Binary128 central differences of every transit output (ntbv = 1: the
    times; 3: times, v_sky, b^2) with respect to the elements:
    ntbv x n x ntt x 7 x n (test_transit_parameters.jl:46-80)."""
    el = f64(elements)
    n = el.shape[0]
    if H is None:
        H = hierarchy_nested(n)
    H = f64(H)
    out = np.zeros((ntbv, n, ntt, 7, n), order="F")
    lib().orc_fd_dtbvdelements(n, C.c_double(t0), dp(el), dp(H), C.c_double(h), C.c_double(tmax), ti,
                               C.c_double(dq0), C.c_longlong(ntt), ntbv, dp(out))
    return out


def set_correction_scale(kicked=1.0, kepler=1.0):
    """Test knob: factors on the 4th-order correction coefficients of phic!
    (kicked pairs) and phisalpha! (Kepler pairs); 1.0 is the reference."""
    lib().orc_set_correction_scale(C.c_double(kicked), C.c_double(kepler))
