"""State, Integrator, TransitTiming, TransitParameters and CartesianOutput: the
reference's user-facing API (src/integrator/Integrator.jl, src/transits/Transits.jl,
src/outputs/Outputs.jl) with the integration running in libnbg_hip.so on the
GPU.

    s = State(ic)
    tt = TransitTiming(tmax, ic)            # ti = 0 (the star) by default
    Integrator(h, tmax)(s, tt, grad=True)   # (::Integrator)(s, tt; grad=true)

Body indices (ti) are 0-based; arrays keep Julia's column-major layout.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import dptr, i64ptr, u8ptr
from .ics import CartesianIC, ElementsIC, init_nbody


class State:
    """Current state of a simulation (Integrator.jl:49-103)."""

    def __init__(self, ic):
        x, v, jac_init = init_nbody(ic)
        n = ic.nbody
        self.n = n
        self.x = x
        self.v = v
        self.t = np.array([float(ic.t0)])
        self.m = ic.m  # shared with the IC, as in the reference
        self.jac_step = np.asfortranarray(np.eye(7 * n))
        self.dqdt = np.zeros(7 * n)
        self.jac_init = jac_init
        self.pair = np.zeros((n, n), dtype=bool, order="F")
        self.xerror = np.zeros((3, n), order="F")
        self.verror = np.zeros((3, n), order="F")
        self.dqdt_error = np.zeros(7 * n)
        self.jac_error = np.zeros((7 * n, 7 * n), order="F")

    def pair_u8(self):
        return np.asfortranarray(self.pair.astype(np.uint8))

    def set_state(self, other):
        """set_state!(self, other) (Integrator.jl:123-134)."""
        self.t[:] = other.t
        self.x[:] = other.x
        self.v[:] = other.v
        self.jac_step[:] = other.jac_step
        self.xerror[:] = other.xerror
        self.verror[:] = other.verror
        self.jac_error[:] = other.jac_error
        self.dqdt[:] = other.dqdt
        self.dqdt_error[:] = other.dqdt_error

    def initialize(self, ic):
        """initialize!(s, ic) (Integrator.jl:112-121)."""
        x, v, jac_init = init_nbody(ic)
        self.x[:] = x
        self.v[:] = v
        self.jac_init[:] = jac_init
        self.xerror[:] = 0.0
        self.verror[:] = 0.0

    def finite(self):
        return bool(np.all(np.isfinite(self.x)) and np.all(np.isfinite(self.v)) and np.all(np.isfinite(self.jac_step)))


class Derivatives:
    """Placeholder for the reference's per-step scratch (PreAllocArrays.jl:65-113):
    the GPU keeps its scratch in LDS, so nothing is allocated here."""

    def __init__(self, T=float, n=0):
        self.n = n


def dState(ic):
    return State(ic), Derivatives(float, ic.nbody)


def _ctx(ctx):
    return ctx if ctx is not None else _lib.default_context()


def ahl21(s, *args, ctx=None):
    """The scheme: ahl21!(s, h) (no Jacobian) or ahl21!(s, d, h) (with
    Jacobian and dq/dt).  One GPU step in place."""
    if len(args) == 1:
        grad, h = False, float(args[0])
    elif len(args) == 2:
        grad, h = True, float(args[1])
    else:
        raise TypeError("ahl21(s, h) or ahl21(s, d, h)")
    c = _ctx(ctx)
    rc = _lib.lib().nbg_ahl21_step(c.handle, s.n, C.c_double(h), int(grad), dptr(s.x), dptr(s.v), dptr(s.xerror),
                                   dptr(s.verror), dptr(s.m), u8ptr(s.pair_u8()), dptr(s.jac_step),
                                   dptr(s.jac_error), dptr(s.dqdt))
    c.check(rc, "nbg_ahl21_step")


class _TransitOutput:
    def _alloc(self, tmax, ic, ti, ntbv):
        if not isinstance(ic, ElementsIC):
            raise TypeError("TransitTiming needs an ElementsIC (Transits.jl:42)")
        n = ic.nbody
        P = ic.elements[:, 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            q = tmax / P
        q = q[np.isfinite(q)]
        self.ntt = int(np.max(np.ceil(np.abs(q)).astype(np.int64) + 3))
        self.ti = int(ti)
        self.n = n
        self.ntbv = ntbv
        self.count = np.zeros(n, dtype=np.int64)
        self.occs = [i for i in range(n) if i != ti]
        self.gsave = np.zeros(n)

    def zero_out(self):
        for a in (self._tt, self._dq, self._del, self.gsave):
            a[...] = 0.0
        self.count[:] = 0


class TransitTiming(_TransitOutput):
    """Transit times and their derivatives (Transits.jl:14-56):
    tt[n, ntt], dtdq0[n, ntt, 7, n], dtdelements[n, ntt, 7, n]."""

    def __init__(self, tmax, ic, ti=0):
        self._alloc(float(tmax), ic, ti, 1)
        n, ntt = self.n, self.ntt
        self.tt = np.zeros((n, ntt), order="F")
        self.dtdq0 = np.zeros((n, ntt, 7, n), order="F")
        self.dtdelements = np.zeros((n, ntt, 7, n), order="F")
        self._tt, self._dq, self._del = self.tt, self.dtdq0, self.dtdelements


class TransitParameters(_TransitOutput):
    """Transit times, sky velocities, impact parameters squared and their
    derivatives (Transits.jl:68-110): ttbv[3, n, ntt], dtbvdq0[3, n, ntt, 7, n]."""

    def __init__(self, tmax, ic, ti=0):
        self._alloc(float(tmax), ic, ti, 3)
        n, ntt = self.n, self.ntt
        self.ttbv = np.zeros((3, n, ntt), order="F")
        self.dtbvdq0 = np.zeros((3, n, ntt, 7, n), order="F")
        self.dtbvdelements = np.zeros((3, n, ntt, 7, n), order="F")
        self._tt, self._dq, self._del = self.ttbv, self.dtbvdq0, self.dtbvdelements


class TransitSnapshot:
    """Impact parameter squared and sky velocity of every body relative to
    body 1 at a list of times, with their derivatives (Transits.jl:112-125):
    vsky, bsky2 [n, nt]; dbvdq0, dbvdelements [2, n, nt, 7, n] (index 0 the
    sky velocity, 1 the impact parameter squared).  The reference's
    constructor builds an undefined TransitSeries (Transits.jl:124); its
    integrator method (Transits.jl:186-211) is what this follows."""

    def __init__(self, times, ic):
        if not isinstance(ic, ElementsIC):
            raise TypeError("TransitSnapshot needs an ElementsIC (Transits.jl:121)")
        self.times = np.asarray(times, dtype=np.float64).copy()
        self.nt = len(self.times)
        n = ic.nbody
        self.n = n
        self.vsky = np.zeros((n, self.nt), order="F")
        self.bsky2 = np.zeros((n, self.nt), order="F")
        self.dbvdq0 = np.zeros((2, n, self.nt, 7, n), order="F")
        self.dbvdelements = np.zeros((2, n, self.nt, 7, n), order="F")


class CartesianOutput:
    """Positions, velocities and Jacobian before every integrator step
    (src/outputs/Outputs.jl:7-21): CartesianOutput(nbody, nstep).  The
    records are arrays over the step index k (last axis): t[nstep],
    x/v/xerror/verror[3, n, nstep], jac_step/jac_error[7n, 7n, nstep]
    (jacobian=False leaves both out).  `states` gives one snapshot per step,
    as the reference's Vector{State}.  With file=True the records are saved
    to `filename` (numpy .npz; the reference writes JLD2)."""

    def __init__(self, nbody, nstep, T=float, filename="data.npz", file=False, jacobian=True):
        if T is not float:
            raise TypeError("only Float64 states are integrated on the GPU")
        n, ns = int(nbody), int(nstep)
        if n < 1 or ns < 0:
            raise ValueError("nbody >= 1 and nstep >= 0")
        self.n = n
        self.nstep = ns
        self.filename = filename
        self.file = bool(file)
        self.t = np.zeros(ns)
        self.x = np.zeros((3, n, ns), order="F")
        self.v = np.zeros((3, n, ns), order="F")
        self.xerror = np.zeros((3, n, ns), order="F")
        self.verror = np.zeros((3, n, ns), order="F")
        self.jacobian = bool(jacobian)
        if self.jacobian:
            self.jac_step = np.zeros((7 * n, 7 * n, ns), order="F")
            self.jac_error = np.zeros((7 * n, 7 * n, ns), order="F")
        else:
            self.jac_step = self.jac_error = None

    def state(self, k):
        """Snapshot before step k (0-based): views into the records."""
        from types import SimpleNamespace
        J = self.jac_step[:, :, k] if self.jacobian else None
        JE = self.jac_error[:, :, k] if self.jacobian else None
        return SimpleNamespace(t=self.t[k], x=self.x[:, :, k], v=self.v[:, :, k], xerror=self.xerror[:, :, k],
                               verror=self.verror[:, :, k], jac_step=J, jac_error=JE)

    @property
    def states(self):
        return [self.state(k) for k in range(self.nstep)]

    def save(self):
        arrays = dict(t=self.t, x=self.x, v=self.v, xerror=self.xerror, verror=self.verror)
        if self.jacobian:
            arrays.update(jac_step=self.jac_step, jac_error=self.jac_error)
        np.savez(self.filename, **arrays)


class Integrator:
    """Integrator(scheme, h, t0, tmax) (Integrator.jl:17-31); callable as
    intr(s), intr(s, time), intr(s, N), intr(s, tt) with grad=True/False."""

    def __init__(self, *args):
        if args and callable(args[0]):
            scheme, rest = args[0], args[1:]
        else:
            scheme, rest = ahl21, args
        if len(rest) == 2:
            h, t0, tmax = rest[0], 0.0, rest[1]
        elif len(rest) == 3:
            h, t0, tmax = rest
        else:
            raise TypeError("Integrator([scheme,] h, [t0,] tmax)")
        if scheme is not ahl21:
            raise ValueError("only the AHL21 scheme is provided by this library")
        self.scheme = scheme
        self.h = float(h)
        self.t0 = float(t0)
        self.tmax = float(tmax)

    def __call__(self, s, arg=None, d=None, grad=True, return_arrays=False, ctx=None):
        if isinstance(arg, Derivatives):
            arg, d = None, arg
        if arg is None:
            return self._to_time(s, s.t[0] + self.tmax, grad, ctx)
        if isinstance(arg, CartesianOutput):
            return self._cartesian(s, arg, grad, ctx)
        if isinstance(arg, TransitSnapshot):
            return self._snapshot(s, arg, grad, ctx)
        if isinstance(arg, _TransitOutput):
            self._transits(s, arg, grad, ctx)
            return Derivatives(float, s.n) if return_arrays else None
        if isinstance(arg, (int, np.integer)) and not isinstance(arg, bool):
            return self._nsteps(s, int(arg), grad, ctx)
        return self._to_time(s, float(arg), grad, ctx)

    # (intr)(s, time; grad): Integrator.jl:159-197
    def _to_time(self, s, time, grad, ctx):
        c = _ctx(ctx)
        t = C.c_double(s.t[0])
        rc = _lib.lib().nbg_integrate_to(c.handle, s.n, C.c_double(self.h), C.c_double(time), int(grad), C.byref(t),
                                         dptr(s.x), dptr(s.v), dptr(s.xerror), dptr(s.verror), dptr(s.m),
                                         u8ptr(s.pair_u8()), dptr(s.jac_step), dptr(s.jac_error), dptr(s.dqdt))
        c.check(rc, "nbg_integrate_to")
        s.t[0] = t.value

    # (intr)(s, N; grad): Integrator.jl:211-234
    def _nsteps(self, s, N, grad, ctx):
        c = _ctx(ctx)
        t = C.c_double(s.t[0])
        rc = _lib.lib().nbg_integrate_steps(c.handle, s.n, C.c_double(self.h), C.c_int64(N), int(grad), C.byref(t),
                                            dptr(s.x), dptr(s.v), dptr(s.xerror), dptr(s.verror), dptr(s.m),
                                            u8ptr(s.pair_u8()), dptr(s.jac_step), dptr(s.jac_error), dptr(s.dqdt))
        c.check(rc, "nbg_integrate_steps")
        s.t[0] = t.value

    # (intr)(s, tt; grad): Transits.jl:140-180
    def _transits(self, s, tt, grad, ctx):
        c = _ctx(ctx)
        t = C.c_double(s.t[0])
        tt.count[:] = 0
        rc = _lib.lib().nbg_integrate_transits(
            c.handle, s.n, C.c_double(self.h), C.c_double(self.tmax), tt.ti, int(grad), tt.ntbv, C.byref(t),
            dptr(s.x), dptr(s.v), dptr(s.xerror), dptr(s.verror), dptr(s.m), u8ptr(s.pair_u8()),
            dptr(s.jac_step), dptr(s.jac_error), dptr(s.dqdt), dptr(s.jac_init), C.c_int64(tt.ntt),
            dptr(tt._tt), dptr(tt._dq), dptr(tt._del), i64ptr(tt.count))
        c.check(rc, "nbg_integrate_transits")
        s.t[0] = t.value

    # (intr)(s, o::CartesianOutput): Outputs.jl:26-49 (the reference always
    # steps with the Jacobian; grad=False records x and v only)
    def _cartesian(self, s, o, grad, ctx):
        if o.n != s.n:
            raise ValueError("CartesianOutput was made for %d bodies, the state has %d" % (o.n, s.n))
        c = _ctx(ctx)
        t = C.c_double(s.t[0])
        jac = o.jacobian and grad
        rc = _lib.lib().nbg_integrate_cartesian(
            c.handle, s.n, C.c_double(self.h), C.c_double(self.tmax), C.c_int64(o.nstep), int(grad), C.byref(t),
            dptr(s.x), dptr(s.v), dptr(s.xerror), dptr(s.verror), dptr(s.m), u8ptr(s.pair_u8()),
            dptr(s.jac_step), dptr(s.jac_error), dptr(s.dqdt), dptr(o.t), dptr(o.x), dptr(o.v), dptr(o.xerror),
            dptr(o.verror), dptr(o.jac_step) if jac else None, dptr(o.jac_error) if jac else None)
        c.check(rc, "nbg_integrate_cartesian")
        s.t[0] = t.value
        if o.file:
            o.save()

    # (intr)(s, ts::TransitSnapshot; grad): Transits.jl:186-211, one call:
    # the GPU integrates to each time and forms b^2, v_sky, their derivatives
    # and the element gradients there
    def _snapshot(self, s, ts, grad, ctx):
        if ts.n != s.n:
            raise ValueError("TransitSnapshot was made for %d bodies, the state has %d" % (ts.n, s.n))
        c = _ctx(ctx)
        t = C.c_double(s.t[0])
        rc = _lib.lib().nbg_integrate_snapshot(
            c.handle, s.n, C.c_double(self.h), C.c_int64(ts.nt), dptr(ts.times), int(grad), C.byref(t),
            dptr(s.x), dptr(s.v), dptr(s.xerror), dptr(s.verror), dptr(s.m), u8ptr(s.pair_u8()),
            dptr(s.jac_step), dptr(s.jac_error), dptr(s.dqdt), dptr(s.jac_init), dptr(ts.vsky), dptr(ts.bsky2),
            dptr(ts.dbvdq0), dptr(ts.dbvdelements))
        c.check(rc, "nbg_integrate_snapshot")
        s.t[0] = t.value
