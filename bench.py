"""Benchmark: AHL21 + Jacobian + transit timing, batched on one MI355X.

Workload (headline): the 4096-system batch -- n ~ U{3..9} bodies, periods
log-uniform in [2, 200] d, h = P_in/30, tmax ~ U[500, 4000] d, FP64 state and
Jacobian, every transit time with its full d t / d elements gradient --
resident in HBM.  One benchmark "step" is one ROUND over the whole batch
(nbg_batch_round): every system advances K integrator steps of its own run
(Transits.jl:140-170: the AHL21 step with Jacobian, the transit search after
every step, calc_dtdelements! when its run ends) in one persistent launch;
systems whose run has ended restart from their initial records, so every
round carries the same mix.  `value` = system-steps (main steps + one per
recorded transit) of the timed rounds / their wall time.  The timed rounds use
the factored Jacobian updates (nbg_set_row_update(1): Jacobians and their
gradients within 1e-10 relative of the reference's order and transit times
within 1e-12, the tolerances asked of them; states bit for bit) unless
--row-update 0; the same rounds in the reference's order are reported beside.

Reported beside it (side legs, each dropped -- never the headline -- when the
time budget would not fit it; they run the library's default, reference-order
row update):
  rounds_reference_order  the headline's rounds in the reference's order;
  c2_makespan  the same batch run once to completion (every system to its tmax);
  plain_gpu    the one-system-per-launch form (nbg_integrate_transits), per-n
               step latency on a sample, summed over the batch's step counts;
  cpu_baseline the CPU restatement (oracle) on a seeded sample of the same
               systems, one system per thread on the host cores; also on one core;
  roofline     algorithmic FP64 flops W(n) = 2016 n^2 (n-1) + 126 n^3 per
               system-step over the dominant launch's HIP-event time, against
               the FP64 vector peak (78.6 TFLOP/s);
  fit_batch    a subset of the 16384-vector fit batch (perturbations of one
               6-body system: star + 5 planets) at its full tmax = 3000 d, initial
               conditions computed on the GPU;
  serving      integrate-plus-gradient requests of the serving trace at full
               tmax: the server's capacity kept full (closed loop), the
               trace's 2000/s arrivals (overload), and arrivals at half the
               capacity (latency p50/p99).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--round-steps S]
                       [--budget SECONDS] [--no-side] ...
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbodygradient.jl_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
# HBM bytes per system-step of the dominant launch in rounds of the 4096-system
# batch, by kernel and row update (0 reference order, 1 factored), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the same counters'
# reading of nbg_copy_kernel of known size (the figures and the commands are
# in profiles/nbg_run_duo_c2_round_counters.md and
# nbg_run_duo_c2_round_factored_counters.md); a configuration without a
# measurement of the current kernels reports traffic null
HBM_BYTES_PER_UNIT = {("nbg_run_duo<9>", 0): 470.3e3, ("nbg_run_duo<9>", 1): 354.7e3}
METRIC = "system-steps/s w/ Jacobian, batched vs plain GPU; p50/p99 request latency"
T_START = time.perf_counter()


def W(n):
    return 2016.0 * n * n * (n - 1) + 126.0 * n ** 3


def log(msg):
    print("[%6.1fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


def plain_latency(specs, states, per_n=2, steps_cap=1500):
    """Per-step latency (s) of the one-system-per-launch form for each n."""
    import nbodygradient_amd as nb
    lat = {}
    for n in sorted(set(s.n for s in specs)):
        idx = [i for i, s in enumerate(specs) if s.n == n][:per_n]
        tot_t, tot_steps = 0.0, 0
        for i in idx:
            sp = specs[i]
            tmax = min(sp.tmax, steps_cap * sp.h)
            ic = nb.ElementsIC(sp.t0, n, sp.elements)
            s = nb.State(ic)
            tt = nb.TransitTiming(tmax, ic)
            t = time.perf_counter()
            nb.Integrator(sp.h, tmax)(s, tt, grad=True)
            tot_t += time.perf_counter() - t
            tot_steps += int(round(tmax / sp.h)) + int(tt.count.sum())
        lat[n] = tot_t / max(tot_steps, 1)
        log("plain n=%d: %.1f us per system-step" % (n, 1e6 * lat[n]))
    return lat


def fit_batch(count, tmax_scale):
    """The fit-style evaluation: `count` Gaussian perturbations of one 6-body
    system (star + 5 planets), every transit time with its dt/delements, the
    initial conditions (x, v, jac_init) computed on the GPU from the elements."""
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.batch import Batch
    specs = Wk.c3(16384)[:count]
    for s in specs:
        s.tmax *= tmax_scale
    t = time.perf_counter()
    b = Batch(specs, device_ic=True)
    t_ic = time.perf_counter() - t
    t = time.perf_counter()
    st = b.run()
    wall = time.perf_counter() - t
    t_run = st["kernel_ms"] / 1e3
    res = b.download(derivatives=False)
    b.close()
    out = {"workload": "fit batch: the first %d of 16384 perturbations of a 6-body system (star + 5 planets, "
                       "P = 5, 8.1, 13.2, 21.6, 35.0 d), h = 0.1 d, tmax %g d, Jacobian + transit times + "
                       "dt/delements, initial conditions on the GPU" % (len(specs), 3000.0 * tmax_scale),
           "systems": len(specs), "system_steps": st["system_steps"], "transits": st["transits"],
           "device_s": t_run, "wall_s": wall,
           "system_steps_per_s": st["system_steps"] / wall, "transits_with_gradients_per_s": st["transits"] / wall,
           "ic_on_gpu_s": t_ic,
           "all_16384_estimate_s": wall * 16384 / len(specs),
           "estimate_basis": "independent systems, throughput-bound batch: time x 16384 / subset",
           "max_count": int(max(int(r["count"].max()) for r in res))}
    log("fit batch: %d systems, %d system-steps in %.2f s (%.3g/s), %d transits; IC on the GPU %.2f s" %
        (len(specs), st["system_steps"], wall, out["system_steps_per_s"], st["transits"], t_ic))
    return out


def cpu_baseline(specs, states, sample, threads, cap_steps):
    """The oracle (CPU restatement) on the first `sample` systems, each capped
    at `cap_steps` steps, one system per OpenMP thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import oracle_lib as O
    from nbodygradient_amd.workloads import SystemSpec
    idx = list(range(min(sample, len(specs))))
    sp = [SystemSpec(specs[i].elements, specs[i].t0, specs[i].h, min(specs[i].tmax, cap_steps * specs[i].h))
          for i in idx]
    st = [states[i] for i in idx]
    n = np.array([s.n for s in sp], dtype=np.int32)
    off_b = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    off_state = (3 * off_b[:-1]).astype(np.int64)
    off_jac = np.concatenate([[0], np.cumsum(49 * n.astype(np.int64) ** 2)])[:-1].astype(np.int64)
    ntt = np.array([s.ntt() for s in sp], dtype=np.int64)
    off_tt = np.concatenate([[0], np.cumsum(n * ntt)])[:-1].astype(np.int64)
    off_dq = np.concatenate([[0], np.cumsum(7 * n.astype(np.int64) ** 2 * ntt)])[:-1].astype(np.int64)
    h = np.array([s.h for s in sp])
    tmax = np.array([s.tmax for s in sp])
    t0 = np.array([s.t0 for s in sp])
    x = np.concatenate([np.asarray(a[0]).ravel(order="F") for a in st])
    v = np.concatenate([np.asarray(a[1]).ravel(order="F") for a in st])
    m = np.concatenate([s.elements[:, 0] for s in sp])
    jinit = np.concatenate([np.asarray(a[2]).ravel(order="F") for a in st])
    tt = np.zeros(int((n * ntt).sum()))
    dq = np.zeros(int((7 * n.astype(np.int64) ** 2 * ntt).sum()))
    de = np.zeros_like(dq)
    count = np.zeros(int(n.sum()), dtype=np.int64)
    done = C.c_longlong(0)
    I32 = C.POINTER(C.c_int)
    secs = O.lib().orc_batch_transits(len(sp), n.ctypes.data_as(I32), O.ip(off_state), O.ip(off_jac), O.ip(off_tt),
                                      O.ip(off_dq), O.dp(h), O.dp(tmax), O.ip(ntt), O.dp(t0), O.dp(x), O.dp(v),
                                      O.dp(m), O.dp(jinit), O.dp(tt), O.dp(dq), O.dp(de), O.ip(count), threads,
                                      C.byref(done))
    return done.value, secs, len(sp)


def serving(trace_count, rate, deadline, chunk, log_fn, label):
    """Requests of the serving trace (n ~ U{2..10}, tmax log-uniform in
    [100, 8000] d, Poisson arrivals) offered at `rate` per second through the
    serving loop, for at most `deadline` seconds."""
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.server import Server, serve
    trace = Wk.c4(count=trace_count, rate=rate)
    arrivals = np.array([t for t, _ in trace])
    sspecs = [sp for _, sp in trace]
    sstates = Wk.initial_states(sspecs)
    sv = Server(max_bytes=64 << 30, chunk=chunk)
    t = time.perf_counter()
    sres = serve(sv, sspecs, sstates, arrivals, log=log_fn, deadline=deadline)
    wall = time.perf_counter() - t
    work = sv.work()
    sv.close()
    lat = np.asarray(sres["latency_s"])
    fin = np.isfinite(lat)
    nfin = int(fin.sum())
    # unfinished requests count as latencies of at least their time in the
    # system: a percentile that falls on one of them is reported as censored
    lat_c = np.where(fin, lat, np.inf)
    srt = np.sort(lat_c)

    def pct(p):
        k = min(len(srt) - 1, int(math.ceil(p / 100.0 * len(srt))) - 1)
        return float(srt[k]) if np.isfinite(srt[k]) else None

    steps_req = np.array([sp.nsteps for sp in sspecs], dtype=np.float64)
    out = {"workload": "%s: %d requests of the serving trace (n ~ U{2..10}, tmax log-uniform in [100, 8000] d, "
                       "h = P_in/30), Poisson arrivals at %g/s, Jacobian + transit times + dt/delements" %
                       (label, len(sspecs), rate),
           "offered_requests_per_s": rate, "chunk_steps": chunk, "deadline_s": deadline, "wall_s": wall,
           "finished": nfin, "unfinished_at_deadline": len(sspecs) - nfin,
           "p50_s": pct(50), "p90_s": pct(90), "p99_s": pct(99),
           "max_finished_s": float(lat[fin].max()) if nfin else None,
           "completed_requests_per_s": nfin / wall,
           "system_steps": work["system_steps"], "system_steps_per_s": work["system_steps"] / wall,
           "device_busy_frac": work["device_ms"] / 1e3 / wall,
           "mean_main_steps_per_request": float(steps_req.mean()), "rounds": work["rounds"]}
    log_fn("serving %s: %d/%d finished in %.1f s, p50 %s, p99 %s, %.3g system-steps/s" %
           (label, nfin, len(sspecs), wall, out["p50_s"], out["p99_s"], out["system_steps_per_s"]))
    return out


def serving_capacity(chunk, log_fn, concurrency=2048, duration=15.0):
    """The server kept full (closed loop, `concurrency` requests of the
    serving trace in flight, each finished one replaced at once) for
    `duration` seconds: its sustained system-step rate, and that rate over the
    trace's mean work per request as requests/s."""
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.server import Server, saturate
    trace = Wk.c4(count=4096)
    sspecs = [sp for _, sp in trace]
    sstates = Wk.initial_states(sspecs)
    sv = Server(max_bytes=64 << 30, chunk=chunk)
    r = saturate(sv, sspecs, sstates, concurrency, duration)
    sv.close()
    per_req = float(np.mean([sp.nsteps + 0.0 for sp in sspecs]))
    rate = r["system_steps"] / r["wall_s"]
    out = {"what": "closed loop: %d requests of the serving trace (n ~ U{2..10}, tmax log-uniform in [100, 8000] d) "
                   "always in flight, each finished one replaced at once, for %g s" % (concurrency, duration),
           "system_steps_per_s": rate, "mean_main_steps_per_request": per_req,
           "requests_per_s": rate / per_req, "completed": r["completed"], "rounds": r["rounds"],
           "device_busy_frac": r["device_ms"] / 1e3 / r["wall_s"], "chunk_steps": chunk}
    log_fn("serving capacity: %.3g system-steps/s = %.1f requests/s (%d completed, %d rounds, device busy %.0f %%)" %
           (rate, out["requests_per_s"], r["completed"], r["rounds"], 100 * out["device_busy_frac"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed rounds")
    ap.add_argument("--warmup", type=int, default=3, help="untimed rounds before them")
    ap.add_argument("--round-steps", type=int, default=500, metavar="K",
                    help="integrator steps every system advances per round")
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--budget", type=float, default=450.0, metavar="SECONDS",
                    help="wall-clock budget of the whole run: a side leg that would not fit is dropped")
    ap.add_argument("--no-side", action="store_true", help="headline only")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--systems-per-wave", type=int, default=0, choices=[0, 1, 2],
                    help="0: the autotuner's choice (1 without it); 1 or 2: forced")
    ap.add_argument("--row-update", type=int, default=1, choices=[0, 1],
                    help="Jacobian updates of the timed rounds: 1 factored (Jacobians within 1e-10 relative of "
                         "the reference's order, transit times within 1e-12, states bitwise), 0 the reference's order")
    ap.add_argument("--reference-rounds", type=int, default=5,
                    help="rounds also timed in the reference's order of the row update (side leg; 0: skip)")
    ap.add_argument("--no-makespan", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--cpu-cap", type=int, default=2000, help="steps per sampled system (all-core run)")
    ap.add_argument("--cpu1-sample", type=int, default=32)
    ap.add_argument("--cpu1-cap", type=int, default=600, help="steps per sampled system (one-core run)")
    ap.add_argument("--fit", type=int, default=1024, metavar="SYSTEMS",
                    help="fit-batch leg: the first SYSTEMS of the 16384 vectors at full tmax (0: skip)")
    ap.add_argument("--fit-tmax-scale", type=float, default=1.0)
    ap.add_argument("--serve", type=int, default=1, help="serving legs (0: skip)")
    ap.add_argument("--serve-chunk", type=int, default=128)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (read MIB, write MIB) through an elementwise "
                         "kernel: the reference reading for the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (systems are independent: a "
              "multi-GPU run would partition the batch across GPUs by estimated cost)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    deadline = T_START + args.budget

    import nbodygradient_amd as nb
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.batch import Batch

    specs = Wk.c2(args.systems)
    states = Wk.initial_states(specs)
    ctx = nb.default_context()
    ctx.set_row_update(args.row_update)
    name, cus = ctx.info()
    log("device %s, %d CUs; %d systems prepared" % (name, cus, len(specs)))

    # prime code loading with a few steps of a tiny batch (not part of the workload)
    pb = Batch(specs[:8], states[:8])
    pb.round(20)
    pb.close()

    K = args.round_steps
    batch = Batch(specs, states, grad=True)
    tune = None
    if not args.no_autotune:
        t = time.perf_counter()
        tune = batch.autotune_round(K)
        c = tune["candidates"][tune["best"]]
        log("autotune (one round of %d steps per candidate, %.1f s): %s ms -> coop mode %d max %d, groups/CU %d" %
            (K, time.perf_counter() - t, ", ".join("%.0f" % q["ms"] for q in tune["candidates"]),
             c["coop_mode"], c["coop_max"], c["groups_per_cu"]) + ", systems per wave %d" % c["systems_per_wave"])
    spw = args.systems_per_wave or (tune["candidates"][tune["best"]]["systems_per_wave"] if tune else 1)
    ctx.set_systems_per_wave(spw)
    for w in range(args.warmup):
        st = batch.round(K)
        log("warmup round %d: %.1f ms device, %d system-steps" % (w, st["kernel_ms"], st["system_steps"]))
    ms, units, trans, newton = [], 0, 0, 0
    launches = {}
    for k in range(args.steps):
        t = time.perf_counter()
        st = batch.round(K)
        wall = time.perf_counter() - t
        ms.append(wall * 1e3)
        units += st["system_steps"]
        trans += st["transits"]
        newton += st["newton_iters"]
        for cls, L in st["launches"].items():
            a = launches.setdefault(cls, {"ms": 0.0, "units": {}, "grid": L["grid"]})
            a["ms"] += L["ms"]
            for n, u in L["units"].items():
                a["units"][n] = a["units"].get(n, 0) + u
        log("round %d: %.1f ms wall, %.1f ms device, %d system-steps, %d transits" %
            (k, wall * 1e3, st["kernel_ms"], st["system_steps"], st["transits"]))
    t_timed = sum(ms) / 1e3
    value = units / t_timed
    ms_per_step = float(np.mean(ms))
    # per-launch roofline over the timed rounds: algorithmic FP64 flops W(n)
    # x units of the launch for n over its duration, from HIP events on the
    # stream it ran on
    lrep = {}
    for cls, L in launches.items():
        flops = sum(W(n) * u for n, u in L["units"].items())
        kname = "nbg_run_coop<10>" if cls == 1 else ("nbg_run_duo<%d>" if spw == 2 else "nbg_run<%d>") % cls
        lrep[str(cls)] = {"kernel": kname, "ms_total": L["ms"], "ms_per_round": L["ms"] / args.steps,
                          "grid": L["grid"], "units": sum(L["units"].values()),
                          "units_by_n": {str(n): u for n, u in sorted(L["units"].items())}, "flops": flops,
                          "tflops": flops / (L["ms"] / 1e3) / 1e12 if L["ms"] > 0 else 0.0}
    dom = max(lrep, key=lambda k: lrep[k]["flops"])
    D = lrep[dom]
    achieved = D["tflops"]
    for c in sorted(lrep, key=int):
        L = lrep[c]
        log("launch %s: grid %d, %.1f ms per round, %d units %s, %.3f TFLOP/s" %
            (L["kernel"], L["grid"], L["ms_per_round"], L["units"], L["units_by_n"], L["tflops"]))
    traffic = None
    bytes_unit = HBM_BYTES_PER_UNIT.get((D["kernel"], args.row_update))
    if bytes_unit is not None:
        traffic = bytes_unit * D["units"] / args.steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "system-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "4096-system batch: n~U{3..9}, P log-U[2,200] d, h=P_in/30, tmax~U[500,4000] d, "
                               "Jacobian + transit times + dt/delements; one step = one round in which every "
                               "system advances %d integrator steps (finished systems restart)" % K,
                   "systems": len(specs), "round_steps": K, "system_steps_timed": units,
                   "transits_timed": trans, "newton_iters_timed": newton,
                   "parallelism": "1 GPU, one persistent launch per body-count class (n <= 9 in one), "
                                  "%d system%s per wave" % (spw, "s" if spw > 1 else ""),
                   "systems_per_wave": spw,
                   "row_update": "factored (Jacobians and their gradients within 1e-10 relative of the reference "
                                 "order, transit times within 1e-12, states bit for bit)" if args.row_update == 1 else
                                 "reference order (every output bit for bit the oracle's)",
                   "autotune": tune},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": D["kernel"], "launch_ms": D["ms_per_round"], "units_per_launch": D["units"] / args.steps,
                     "note": "FP64 vector (VALU) peak 78.6 TFLOP/s; no MFMA in the kernel (a measured pair-update "
                             "probe ran 28 % slower on the FP64 matrix cores); W(n) = 2016 n^2 (n-1) + 126 n^3 flops "
                             "per system-step, the reference formulation's chain work whatever the updates' form "
                             "(the factored form needs fewer operations for it); traffic = " +
                             ("%.1f KB per system-step (calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of these rounds) "
                              "x the launch's system-steps per round" % (bytes_unit / 1e3) if bytes_unit else
                              "not measured for this kernel")},
        "launches": lrep,
        "transits_with_gradients_per_s": trans / t_timed,
    }
    log("headline: %.4g system-steps/s (%.1f ms per round), %s at %.3f TFLOP/s = %.4f of peak" %
        (value, ms_per_step, D["kernel"], achieved, achieved / FP64_PEAK_TFLOPS))
    skipped = []

    def fits(est, leg):
        if args.no_side:
            return False
        if time.perf_counter() + est > deadline:
            skipped.append(leg)
            log("side leg %s dropped: %.0f s left, needs ~%.0f s" % (leg, deadline - time.perf_counter(), est))
            return False
        return True

    if args.row_update == 1 and args.reference_rounds > 0 and fits(8, "rounds_reference_order"):
        # the same rounds with the row update in the reference's order (bit for bit the oracle's results)
        ctx.set_row_update(0)
        batch.round(K)
        rms, runits = [], 0
        for k in range(args.reference_rounds):
            t = time.perf_counter()
            st = batch.round(K)
            rms.append((time.perf_counter() - t) * 1e3)
            runits += st["system_steps"]
        out["rounds_reference_order"] = {"value": runits / (sum(rms) / 1e3), "unit": "system-steps/s",
                                         "ms_per_round": float(np.mean(rms)), "rounds": args.reference_rounds,
                                         "what": "the headline's rounds with the Jacobian row update in the reference's "
                                                 "order: every output bit for bit the oracle's"}
        log("rounds, reference order of the row update: %.4g system-steps/s (%.1f ms per round)" %
            (out["rounds_reference_order"]["value"], np.mean(rms)))
    ctx.set_row_update(0)  # the side legs below run the library's default form
    if not args.no_makespan and fits(45, "c2_makespan"):
        ctx.set_coop(1, 128)  # whole runs: the automatic two-wave selection, four one-wave groups per CU
        ctx.set_launch(0)
        batch.reset()
        t = time.perf_counter()
        st = batch.run()
        wall = time.perf_counter() - t
        out["c2_makespan"] = {"wall_s": wall, "device_s": st["kernel_ms"] / 1e3, "system_steps": st["system_steps"],
                              "transits": st["transits"], "system_steps_per_s": st["system_steps"] / wall,
                              "what": "the 4096-system batch once to completion (every system to its tmax), "
                                      "the longest systems on two-wave workgroups (automatic)"}
        log("makespan: %.2f s wall, %d system-steps, %d transits" % (wall, st["system_steps"], st["transits"]))
    plain_rate = None
    if not args.no_plain and fits(15, "plain_gpu"):
        lat = plain_latency(specs, states)
        # the batch's work in whole runs, priced at the one-system form's per-step latency
        t_plain, tot = 0.0, 0
        for sp in specs:
            t_plain += lat[sp.n] * sp.nsteps
            tot += sp.nsteps
        plain_rate = tot / t_plain
        out["plain_gpu"] = {"value": plain_rate, "unit": "system-steps/s",
                            "method": "one system per call (nbg_integrate_transits, host arrays in and out); per-n "
                                      "step latency from 2 systems x <= 1500 steps, priced over the batch's main steps"}
        out["batched_vs_plain"] = value / plain_rate
    if not args.no_cpu and fits(40, "cpu_baseline"):
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        done, secs, ns = cpu_baseline(specs, states, args.cpu_sample, threads, args.cpu_cap)
        cpu_rate = done / secs
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "system-steps/s", "cores": threads, "kind": "port",
                               "sample": "first %d of the %d systems (seeded), each capped at %d steps: %d system-steps "
                                         "in %.1f s, OpenMP one system per thread" %
                                         (ns, len(specs), args.cpu_cap, done, secs)}
        out["batched_vs_cpu"] = value / cpu_rate
        log("cpu: %d threads, %.4g system-steps/s" % (threads, cpu_rate))
        done1, secs1, ns1 = cpu_baseline(specs, states, args.cpu1_sample, 1, args.cpu1_cap)
        out["cpu_baseline_1core"] = {"value": done1 / secs1, "unit": "system-steps/s", "cores": 1, "kind": "port",
                                     "sample": "first %d systems, each capped at %d steps: %d system-steps in %.1f s" %
                                               (ns1, args.cpu1_cap, done1, secs1)}
        log("cpu 1 core: %.4g system-steps/s" % (done1 / secs1))
    if args.fit > 0 and fits(40, "fit_batch"):
        out["fit_batch"] = fit_batch(args.fit, args.fit_tmax_scale)
    if args.serve > 0 and fits(30, "serving_capacity"):
        ctx.set_coop(1, 128)  # serving rounds keep the defaults, whatever the autotuner chose
        ctx.set_launch(0)
        out["serving_capacity"] = serving_capacity(args.serve_chunk, log)
    if args.serve > 0 and fits(35, "serving_overload"):
        # the trace's own rate, 2000 requests/s of full-tmax requests: several
        # times the capacity, so requests queue; bounded at 25 s
        out["serving_overload"] = serving(2000, 2000.0, 25.0, args.serve_chunk, log, "overload at 2000/s")
    if args.serve > 0 and "serving_capacity" in out and fits(110, "serving_latency"):
        # latency at half the measured capacity: 12 s of arrivals, drained
        cap = out["serving_capacity"]["requests_per_s"]
        rate = max(1.0, round(0.5 * cap))
        nreq = int(min(4000, max(50, rate * 12.0)))
        out["serving"] = serving(nreq, rate, 100.0, args.serve_chunk, log, "latency at half capacity")
    out["side_legs_dropped"] = skipped
    out["bench_wall_s"] = time.perf_counter() - T_START
    if args.calibrate_copy > 0:
        import ctypes as C
        from nbodygradient_amd import _lib
        nbytes = args.calibrate_copy * (1 << 20)
        cms = C.c_double(0.0)
        ctx.check(_lib.lib().nbg_copy_calibrate(ctx.handle, nbytes, 3, C.byref(cms)), "nbg_copy_calibrate")
        out["copy_calibration"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": cms.value,
                                   "GB_per_s": 2 * nbytes / (cms.value / 1e3) / 1e9, "kernel": "nbg_copy_kernel"}
        log("copy calibration: nbg_copy_kernel, %d MiB read + %d MiB written, %.3f ms, %.0f GB/s" %
            (args.calibrate_copy, args.calibrate_copy, cms.value, 2 * nbytes / (cms.value / 1e3) / 1e9))
    batch.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
