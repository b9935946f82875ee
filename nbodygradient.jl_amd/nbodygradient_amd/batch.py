"""Batched form: many independent systems integrated in one persistent GPU
launch per body count (C ABI: nbg_batch_create / run / download)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BatchStats, dptr, i32ptr, i64ptr, u8ptr
from .workloads import initial_states


class Batch:
    """A set of systems resident on the GPU.

    specs: list of SystemSpec (elements, t0, h, tmax); states: optional list
    of (x, v, jac_init) as returned by init_nbody (computed on the host when
    omitted, or on the GPU with device_ic=True: fully nested hierarchies).
    Each system runs (intr)(s, tt; grad) with TransitTiming (ntbv = 1) or
    TransitParameters (ntbv = 3), transits measured against body ti."""

    def __init__(self, specs, states=None, grad=True, ntbv=1, ti=0, pair=None, ctx=None, device_ic=False):
        self.ctx = ctx if ctx is not None else _lib.default_context()
        self.specs = list(specs)
        if device_ic:
            if states is not None:
                raise ValueError("device_ic computes the states; pass one or the other")
            self._create_elements(grad, ntbv, ti, pair)
            return
        if states is None:
            states = initial_states(self.specs)
        ns = len(self.specs)
        self._arrays(grad, ntbv, ti)
        x = np.concatenate([np.asarray(st[0], order="F").ravel(order="F") for st in states])
        v = np.concatenate([np.asarray(st[1], order="F").ravel(order="F") for st in states])
        m = np.concatenate([s.elements[:, 0] for s in self.specs])
        jinit = np.concatenate([np.asarray(st[2], order="F").ravel(order="F") for st in states])
        pr = None
        if pair is not None:
            pr = np.concatenate([np.asarray(p, dtype=np.uint8).ravel(order="F") for p in pair])
        self._x, self._v, self._m, self._jinit, self._pair = x, v, m, jinit, pr
        h = C.c_void_p()
        rc = _lib.lib().nbg_batch_create(self.ctx.handle, ns, i32ptr(self.n), dptr(self.h), dptr(self.tmax),
                                         dptr(self.t0), i32ptr(self.ti), i64ptr(self.ntt), int(grad), self.ntbv,
                                         dptr(x), dptr(v), dptr(m), u8ptr(pr), dptr(jinit), C.byref(h))
        self.ctx.check(rc, "nbg_batch_create")
        self._h = h
        self._offsets()

    def _arrays(self, grad, ntbv, ti):
        ns = len(self.specs)
        self.n = np.array([s.n for s in self.specs], dtype=np.int32)
        self.h = np.array([s.h for s in self.specs], dtype=np.float64)
        self.tmax = np.array([s.tmax for s in self.specs], dtype=np.float64)
        self.t0 = np.array([s.t0 for s in self.specs], dtype=np.float64)
        self.ti = np.full(ns, ti, dtype=np.int32)
        self.ntt = np.array([s.ntt() for s in self.specs], dtype=np.int64)
        self.grad = bool(grad)
        self.ntbv = int(ntbv)

    def _create_elements(self, grad, ntbv, ti, pair):
        """nbg_batch_create_elements: init_nbody on the GPU."""
        self._arrays(grad, ntbv, ti)
        el = np.concatenate([np.asfortranarray(s.elements[:, :7], dtype=np.float64).ravel(order="F")
                             for s in self.specs])
        pr = None
        if pair is not None:
            pr = np.concatenate([np.asarray(p, dtype=np.uint8).ravel(order="F") for p in pair])
        self._el, self._pair = el, pr
        h = C.c_void_p()
        rc = _lib.lib().nbg_batch_create_elements(self.ctx.handle, len(self.specs), i32ptr(self.n), dptr(self.h),
                                                  dptr(self.tmax), dptr(self.t0), i32ptr(self.ti), i64ptr(self.ntt),
                                                  int(grad), self.ntbv, dptr(el), None, u8ptr(pr), C.byref(h))
        self.ctx.check(rc, "nbg_batch_create_elements")
        self._h = h
        self._offsets()

    def initial(self):
        """The batch's initial (x, v, jac_init) per system."""
        nbod = self.n.astype(np.int64)
        x = np.zeros(3 * int(nbod.sum()))
        v = np.zeros(3 * int(nbod.sum()))
        J = np.zeros(int((49 * nbod ** 2).sum()))
        self.ctx.check(_lib.lib().nbg_batch_initial(self._h, dptr(x), dptr(v), dptr(J)), "nbg_batch_initial")
        out, ob, oj = [], 0, 0
        for n in nbod:
            n = int(n)
            out.append((x[3 * ob:3 * (ob + n)].reshape((3, n), order="F"),
                        v[3 * ob:3 * (ob + n)].reshape((3, n), order="F"),
                        J[oj:oj + 49 * n * n].reshape((7 * n, 7 * n), order="F")))
            ob += n
            oj += 49 * n * n
        return out

    def autotune(self, frac=0.05):
        """nbg_batch_autotune: time a copy of this batch with every step count
        scaled by frac under each candidate launch configuration (two-wave
        workgroups off / automatic / for the 16, 32, 64 longest systems;
        four or three one-wave workgroups per CU; one or two systems per
        wave) and keep the fastest in the context.  Returns the candidates with their times and the choice."""
        L = _lib.lib()
        nc = L.nbg_batch_autotune_candidate(-1, None, None, None)
        ms = np.zeros(nc)
        best = C.c_int32(-1)
        self.ctx.check(L.nbg_batch_autotune(self._h, C.c_double(frac), C.byref(best), dptr(ms), nc),
                       "nbg_batch_autotune")
        cands = []
        for k in range(nc):
            a, b_, g = C.c_int32(), C.c_int32(), C.c_int32()
            L.nbg_batch_autotune_candidate(k, C.byref(a), C.byref(b_), C.byref(g))
            cands.append({"coop_mode": a.value, "coop_max": b_.value, "groups_per_cu": g.value,
                          "systems_per_wave": L.nbg_batch_autotune_systems_per_wave(k), "ms": float(ms[k])})
        return {"frac": frac, "best": best.value, "candidates": cands}

    def autotune_round(self, steps):
        """nbg_batch_autotune_round: time one round of `steps` steps from the
        initial state under each candidate launch configuration, keep the
        fastest in the context and reset the batch."""
        L = _lib.lib()
        nc = L.nbg_batch_autotune_candidate(-1, None, None, None)
        ms = np.zeros(nc)
        best = C.c_int32(-1)
        self.ctx.check(L.nbg_batch_autotune_round(self._h, C.c_int64(steps), C.byref(best), dptr(ms), nc),
                       "nbg_batch_autotune_round")
        cands = []
        for k in range(nc):
            a, b_, g = C.c_int32(), C.c_int32(), C.c_int32()
            L.nbg_batch_autotune_candidate(k, C.byref(a), C.byref(b_), C.byref(g))
            cands.append({"coop_mode": a.value, "coop_max": b_.value, "groups_per_cu": g.value,
                          "systems_per_wave": L.nbg_batch_autotune_systems_per_wave(k), "ms": float(ms[k])})
        return {"round_steps": int(steps), "best": best.value, "candidates": cands}

    def _offsets(self):
        nb = self.ntbv
        self._off_tt = np.concatenate([[0], np.cumsum(nb * self.n.astype(np.int64) * self.ntt)])
        self._off_dq = np.concatenate([[0], np.cumsum(nb * self.n.astype(np.int64) ** 2 * 7 * self.ntt)])
        self._off_b = np.concatenate([[0], np.cumsum(self.n.astype(np.int64))])

    @property
    def main_steps(self):
        return int(sum(s.nsteps for s in self.specs))

    @staticmethod
    def _stats(st):
        return {"system_steps": st.system_steps, "transits": st.transits, "newton_iters": st.newton_iters,
                "kernel_ms": st.kernel_ms, "paired_systems": st.paired_systems,
                "launches": {c: {"ms": st.launch_ms[c], "grid": st.launch_grid[c],
                                 "units": {n: st.launch_units[c][n] for n in range(13) if st.launch_units[c][n] > 0}}
                             for c in range(13) if st.launch_ms[c] > 0}}

    def run(self):
        """Integrate every system to the end of its run (nbg_batch_run)."""
        st = BatchStats()
        self.ctx.check(_lib.lib().nbg_batch_run(self._h, C.byref(st)), "nbg_batch_run")
        return self._stats(st)

    def round(self, steps, restart=True):
        """One round (nbg_batch_round): every system advances at most `steps`
        steps of its run; with restart, systems whose run had ended start it
        again from their initial records first.  The stats are this round's."""
        st = BatchStats()
        self.ctx.check(_lib.lib().nbg_batch_round(self._h, C.c_int64(steps), int(bool(restart)), C.byref(st)),
                       "nbg_batch_round")
        return self._stats(st)

    def reset(self):
        self.ctx.check(_lib.lib().nbg_batch_reset(self._h), "nbg_batch_reset")

    def download(self, derivatives=True):
        """Per-system results: list of dicts with tt, dtdq0, dtdelements, count, x, v."""
        tt = np.zeros(int(self._off_tt[-1]))
        dq = np.zeros(int(self._off_dq[-1])) if derivatives else None
        de = np.zeros(int(self._off_dq[-1])) if derivatives else None
        count = np.zeros(int(self._off_b[-1]), dtype=np.int64)
        x = np.zeros(3 * int(self._off_b[-1]))
        v = np.zeros(3 * int(self._off_b[-1]))
        rc = _lib.lib().nbg_batch_download(self._h, dptr(tt), dptr(dq), dptr(de), i64ptr(count), dptr(x), dptr(v))
        self.ctx.check(rc, "nbg_batch_download")
        out = []
        nb = self.ntbv
        for s in range(len(self.specs)):
            n, ntt = int(self.n[s]), int(self.ntt[s])
            shp_tt = (n, ntt) if nb == 1 else (3, n, ntt)
            shp_dq = (n, ntt, 7, n) if nb == 1 else (3, n, ntt, 7, n)
            r = {"tt": tt[self._off_tt[s]:self._off_tt[s + 1]].reshape(shp_tt, order="F"),
                 "count": count[self._off_b[s]:self._off_b[s + 1]].copy(),
                 "x": x[3 * self._off_b[s]:3 * self._off_b[s + 1]].reshape((3, n), order="F"),
                 "v": v[3 * self._off_b[s]:3 * self._off_b[s + 1]].reshape((3, n), order="F")}
            if derivatives:
                r["dtdq0"] = dq[self._off_dq[s]:self._off_dq[s + 1]].reshape(shp_dq, order="F")
                r["dtdelements"] = de[self._off_dq[s]:self._off_dq[s + 1]].reshape(shp_dq, order="F")
            out.append(r)
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().nbg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def snapshot_batch(specs, times, states=None, grad=True, pair=None, ctx=None):
    """TransitSnapshot for many systems in one launch (nbg_batch_snapshot):
    system s starts at specs[s].t0 from states[s] = (x, v, jac_init) (host
    init_nbody when omitted), steps with specs[s].h and is integrated to each
    of times[s].  Returns per system a dict with vsky, bsky2 [n, nt] and, with
    grad, dbvdq0, dbvdelements [2, n, nt, 7, n]."""
    ctx = ctx if ctx is not None else _lib.default_context()
    specs = list(specs)
    if states is None:
        states = initial_states(specs)
    ns = len(specs)
    n = np.array([s.n for s in specs], dtype=np.int32)
    h = np.array([s.h for s in specs], dtype=np.float64)
    t0 = np.array([s.t0 for s in specs], dtype=np.float64)
    tl = [np.asarray(t, dtype=np.float64) for t in times]
    nt = np.array([len(t) for t in tl], dtype=np.int64)
    tt = np.concatenate(tl)
    x = np.concatenate([np.asarray(st[0], order="F").ravel(order="F") for st in states])
    v = np.concatenate([np.asarray(st[1], order="F").ravel(order="F") for st in states])
    m = np.concatenate([s.elements[:, 0] for s in specs])
    ji = np.concatenate([np.asarray(st[2], order="F").ravel(order="F") for st in states])
    pr = None
    if pair is not None:
        pr = np.concatenate([np.asarray(p, dtype=np.uint8).ravel(order="F") for p in pair])
    n64 = n.astype(np.int64)
    vs = np.zeros(int((n64 * nt).sum()))
    bs = np.zeros(int((n64 * nt).sum()))
    dq = np.zeros(int((14 * n64 ** 2 * nt).sum())) if grad else None
    de = np.zeros(int((14 * n64 ** 2 * nt).sum())) if grad else None
    rc = _lib.lib().nbg_batch_snapshot(ctx.handle, ns, i32ptr(n), dptr(h), dptr(t0), i64ptr(nt), dptr(tt), int(grad),
                                       dptr(x), dptr(v), dptr(m), u8ptr(pr), dptr(ji), dptr(vs), dptr(bs), dptr(dq),
                                       dptr(de))
    ctx.check(rc, "nbg_batch_snapshot")
    out, oo, oq = [], 0, 0
    for k, c in zip(n64, nt):
        k, c = int(k), int(c)
        r = {"vsky": vs[oo:oo + k * c].reshape((k, c), order="F"),
             "bsky2": bs[oo:oo + k * c].reshape((k, c), order="F")}
        if grad:
            r["dbvdq0"] = dq[oq:oq + 14 * k * k * c].reshape((2, k, c, 7, k), order="F")
            r["dbvdelements"] = de[oq:oq + 14 * k * k * c].reshape((2, k, c, 7, k), order="F")
        out.append(r)
        oo += k * c
        oq += 14 * k * k * c
    return out


def cartesian_batch(specs, nstep, states=None, grad=True, errors=True, jacobian=False, pair=None, ctx=None):
    """CartesianOutput for many systems in one launch (nbg_batch_cartesian):
    system s takes nstep[s] steps of specs[s].h from specs[s].t0 towards
    specs[s].tmax, starting from states[s] = (x, v, jac_init) (host init_nbody
    when omitted), and records the state before each step.  Returns per
    system a dict of t [nstep], x, v (and xerror, verror) [3, n, nstep], and
    with jacobian=True jac_step, jac_error [7n, 7n, nstep]."""
    ctx = ctx if ctx is not None else _lib.default_context()
    specs = list(specs)
    if states is None:
        states = initial_states(specs)
    ns = len(specs)
    n = np.array([s.n for s in specs], dtype=np.int32)
    h = np.array([s.h for s in specs], dtype=np.float64)
    tmax = np.array([s.tmax for s in specs], dtype=np.float64)
    t0 = np.array([s.t0 for s in specs], dtype=np.float64)
    nst = np.asarray(nstep, dtype=np.int64)
    x = np.concatenate([np.asarray(st[0], order="F").ravel(order="F") for st in states])
    v = np.concatenate([np.asarray(st[1], order="F").ravel(order="F") for st in states])
    m = np.concatenate([s.elements[:, 0] for s in specs])
    pr = None
    if pair is not None:
        pr = np.concatenate([np.asarray(p, dtype=np.uint8).ravel(order="F") for p in pair])
    flags = (1 if errors else 0) | (6 if jacobian else 0)
    n64 = n.astype(np.int64)
    T = np.zeros(int(nst.sum()))
    X = np.zeros(int((3 * n64 * nst).sum()))
    V = np.zeros_like(X)
    XE = np.zeros_like(X) if errors else None
    VE = np.zeros_like(X) if errors else None
    J = np.zeros(int((49 * n64 ** 2 * nst).sum())) if jacobian else None
    JE = np.zeros_like(J) if jacobian else None
    rc = _lib.lib().nbg_batch_cartesian(ctx.handle, ns, i32ptr(n), dptr(h), dptr(tmax), dptr(t0), i64ptr(nst),
                                        int(grad), flags, dptr(x), dptr(v), dptr(m), u8ptr(pr), dptr(T), dptr(X),
                                        dptr(V), dptr(XE), dptr(VE), dptr(J), dptr(JE))
    ctx.check(rc, "nbg_batch_cartesian")
    out, o1, o3, oj = [], 0, 0, 0
    for k, c in zip(n64, nst):
        k, c = int(k), int(c)
        r = {"t": T[o1:o1 + c], "x": X[o3:o3 + 3 * k * c].reshape((3, k, c), order="F"),
             "v": V[o3:o3 + 3 * k * c].reshape((3, k, c), order="F")}
        if errors:
            r["xerror"] = XE[o3:o3 + 3 * k * c].reshape((3, k, c), order="F")
            r["verror"] = VE[o3:o3 + 3 * k * c].reshape((3, k, c), order="F")
        if jacobian:
            r["jac_step"] = J[oj:oj + 49 * k * k * c].reshape((7 * k, 7 * k, c), order="F")
            r["jac_error"] = JE[oj:oj + 49 * k * k * c].reshape((7 * k, 7 * k, c), order="F")
        out.append(r)
        o1 += c
        o3 += 3 * k * c
        oj += 49 * k * k * c
    return out
