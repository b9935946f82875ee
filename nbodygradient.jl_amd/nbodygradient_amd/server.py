"""Serving: continuous batching of integrate-plus-gradient requests on one GPU
(C ABI: nbg_server_*).

A Server owns a device arena.  submit() queues a request (one system's
(intr)(s, tt; grad) run); step() admits queued requests while their HBM
footprint fits, advances every admitted request by up to `chunk` steps in one
set of persistent launches and retires the finished ones; fetch() returns a
finished request's TransitTiming outputs.  serve() drives a server in real
time from a trace of arrival times and reports per-request latency.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import _lib
from ._lib import dptr, i64ptr, u8ptr


class Server:
    def __init__(self, max_bytes=8 << 30, chunk=256, ctx=None):
        self.ctx = ctx if ctx is not None else _lib.default_context()
        h = C.c_void_p()
        self.ctx.check(_lib.lib().nbg_server_create(self.ctx.handle, int(max_bytes), int(chunk), C.byref(h)),
                       "nbg_server_create")
        self._h = h
        self._meta = {}
        self._lock = threading.Lock()

    @staticmethod
    def footprint(n, ntt, ntbv=1):
        return int(_lib.lib().nbg_server_footprint(int(n), int(ntbv), int(ntt)))

    def submit(self, spec, state, ti=0, grad=True, ntbv=1, pair=None):
        """Queue (intr)(s, tt; grad) for one system; spec: SystemSpec, state:
        (x, v, jac_init) from init_nbody.  Returns the request id."""
        x = np.asarray(state[0], dtype=np.float64, order="F")
        v = np.asarray(state[1], dtype=np.float64, order="F")
        jinit = np.asarray(state[2], dtype=np.float64, order="F")
        m = np.ascontiguousarray(spec.elements[:, 0], dtype=np.float64)
        pr = None if pair is None else np.asarray(pair, dtype=np.uint8, order="F")
        rid = C.c_int64(0)
        rc = _lib.lib().nbg_server_submit(self._h, spec.n, spec.h, spec.tmax, spec.t0, int(ti), spec.ntt(), int(grad),
                                          int(ntbv), dptr(x), dptr(v), dptr(m), u8ptr(pr), dptr(jinit), C.byref(rid))
        self.ctx.check(rc, "nbg_server_submit")
        with self._lock:
            self._meta[rid.value] = (spec.n, spec.ntt(), int(ntbv), bool(grad))
        return rid.value

    def step(self):
        """One round; returns (number finished, device ms)."""
        nf = C.c_int32(0)
        ms = C.c_double(0.0)
        self.ctx.check(_lib.lib().nbg_server_step(self._h, C.byref(nf), C.byref(ms)), "nbg_server_step")
        return nf.value, ms.value

    def finished(self, max_ids=4096):
        ids = np.zeros(max_ids, dtype=np.int64)
        n = C.c_int32(0)
        self.ctx.check(_lib.lib().nbg_server_finished(self._h, i64ptr(ids), int(max_ids), C.byref(n)),
                       "nbg_server_finished")
        return [int(i) for i in ids[:n.value]]

    def fetch(self, rid):
        with self._lock:
            n, ntt, nb, grad = self._meta.pop(rid)
        shp_tt = (n, ntt) if nb == 1 else (3, n, ntt)
        shp_dq = (n, ntt, 7, n) if nb == 1 else (3, n, ntt, 7, n)
        tt = np.zeros(shp_tt, order="F")
        dq = np.zeros(shp_dq, order="F") if grad else None
        de = np.zeros(shp_dq, order="F") if grad else None
        count = np.zeros(n, dtype=np.int64)
        x = np.zeros((3, n), order="F")
        v = np.zeros((3, n), order="F")
        rc = _lib.lib().nbg_server_fetch(self._h, int(rid), dptr(tt), dptr(dq), dptr(de), i64ptr(count), dptr(x),
                                         dptr(v))
        self.ctx.check(rc, "nbg_server_fetch")
        return {"tt": tt, "dtdq0": dq, "dtdelements": de, "count": count, "x": x, "v": v}

    def status(self):
        vals = [C.c_int64(0) for _ in range(4)]
        self.ctx.check(_lib.lib().nbg_server_status(self._h, *[C.byref(q) for q in vals]), "nbg_server_status")
        return dict(zip(["pending", "active", "finished", "bytes_free"], [q.value for q in vals]))

    def work(self):
        """System-steps (main steps + one per recorded transit), rounds and
        device ms over every round so far."""
        st, rd, ms = C.c_int64(0), C.c_int64(0), C.c_double(0.0)
        self.ctx.check(_lib.lib().nbg_server_work(self._h, C.byref(st), C.byref(rd), C.byref(ms)), "nbg_server_work")
        return {"system_steps": st.value, "rounds": rd.value, "device_ms": ms.value}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().nbg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(server, specs, states, arrivals, keep_results=False, log=None, deadline=None):
    """Drive `server` in real time: request k is submitted at arrivals[k]
    seconds after the start (a separate thread), rounds run back to back on
    the calling thread.  Returns per-request latency (s, arrival to result
    fetched; NaN for a request still unfinished when `deadline` seconds
    have passed, where the loop stops), the wall time and, optionally, the
    results by request index."""
    order = np.argsort(arrivals, kind="stable")
    n = len(specs)
    submit_t = np.zeros(n)
    done_t = np.full(n, np.nan)
    ids = {}
    t0 = time.perf_counter()
    submitted = threading.Event()
    maplock = threading.Lock()  # a request may finish before submit() has returned to the producer
    errors = []

    def producer():
        try:
            for k in order:
                delay = arrivals[k] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                with maplock:
                    submit_t[k] = time.perf_counter() - t0
                    rid = server.submit(specs[k], states[k])
                    ids[rid] = k
        except Exception as e:  # surfaced by the caller
            errors.append(e)
        finally:
            submitted.set()

    th = threading.Thread(target=producer, daemon=True)
    th.start()
    results = {}
    ndone = 0
    rounds = 0
    device_ms = 0.0
    while ndone < n and not errors:
        if deadline is not None and time.perf_counter() - t0 > deadline and submitted.is_set():
            break
        nf, ms = server.step()
        rounds += 1
        device_ms += ms
        if nf == 0 and ms == 0.0:
            time.sleep(0.0005)  # nothing admitted yet
        for rid in server.finished():
            r = server.fetch(rid)
            with maplock:
                k = ids[rid]
            done_t[k] = time.perf_counter() - t0
            ndone += 1
            if keep_results:
                results[k] = r
        if log and rounds % 50 == 0:
            log("serve: round %d, %d/%d done, %s" % (rounds, ndone, n, server.status()))
    th.join()
    if errors:
        raise errors[0]
    wall = time.perf_counter() - t0
    lat = done_t - submit_t
    return {"latency_s": lat, "wall_s": wall, "rounds": rounds, "device_ms": device_ms, "results": results,
            "finished": ndone}


def saturate(server, specs, states, concurrency, duration, log=None):
    """Closed loop: keep `concurrency` requests in the server (each finished
    request is replaced at once by the next of specs, cycling) for `duration`
    seconds of rounds, so that the GPU never waits for arrivals.  Returns the
    system-steps (main steps + one per recorded transit), requests completed
    and rounds over that time: the server's sustained capacity."""
    n = len(specs)
    nxt = 0
    live = 0
    while live < concurrency:
        server.submit(specs[nxt % n], states[nxt % n])
        nxt += 1
        live += 1
    w0 = server.work()
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < duration:
        server.step()
        for rid in server.finished():
            server.fetch(rid)
            done += 1
            server.submit(specs[nxt % n], states[nxt % n])
            nxt += 1
        if log and server.work()["rounds"] % 50 == 0:
            log("saturate: %d completed, %s" % (done, server.status()))
    wall = time.perf_counter() - t0
    w1 = server.work()
    return {"wall_s": wall, "system_steps": w1["system_steps"] - w0["system_steps"], "completed": done,
            "rounds": w1["rounds"] - w0["rounds"], "device_ms": w1["device_ms"] - w0["device_ms"]}

