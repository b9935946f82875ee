"""Benchmark: AHL21 + Jacobian + transit timing, batched on one MI355X.

Workload (default): 4096 independent systems, n ~ U{3..9} bodies, periods
log-uniform in [2, 200] d, h = P_in/30, tmax ~ U[500, 4000] d, FP64 state and
Jacobian, every transit time with its full d t / d elements gradient.  A
"step" of this benchmark is one complete pass of the hot path over the whole
batch: (intr)(s, tt; grad=true) for all 4096 systems in one persistent launch
per body count, inputs already resident in HBM.

Reported beside it:
  plain_gpu    the one-system-per-launch form on a stratified sample, per-step
               latency measured per n and summed over the batch's step counts;
  cpu_baseline the CPU restatement (oracle) on a seeded sample of the same
               systems, one system per thread on all host cores;
  roofline     algorithmic FP64 flops W(n) = 2016 n^2 (n-1) + 126 n^3 per
               system-step (the structure-aware Jacobian-chain work), over the
               batch's device time, against the FP64 peak (78.6 TFLOP/s);
  serving      the C4 request trace through the serving loop (p50/p99);
  fit_batch    the C3 fit-style batch (16384 perturbed 5-body systems, tmax
               scaled), initial conditions computed on the GPU from the
               elements, beside the host's init_nbody for the same systems.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--systems N]
                       [--tmax-scale S] [--no-plain] [--no-cpu] [--serve R]
                       [--fit SYSTEMS] [--autotune FRAC] [--coop auto|off|K]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbodygradient.jl_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
# HBM bytes per system-step of nbg_run<9> (every n <= 9 in one launch), from
# the rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the same counters'
# reading of a copy of known size (profiles/nbg_run_c2_x0.05_counters.md:
# 700.9 GB read + 914.1 GB written for the 4,174,120 system-steps nbg_run<9>
# carries of the C2 x0.05 batch, in one dispatch)
HBM_BYTES_PER_UNIT = (700.9e9 + 914.1e9) / 4174120
METRIC = "system-steps/s w/ Jacobian, batched vs plain GPU; p50/p99 request latency"


def W(n):
    return 2016.0 * n * n * (n - 1) + 126.0 * n ** 3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    """The C3 fit-style evaluation: `count` Gaussian perturbations of one
    5-body system, every transit time with its dt/delements, the initial
    conditions (x, v, jac_init) computed on the GPU from the elements."""
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.batch import Batch
    specs = Wk.c3(count)
    for s in specs:
        s.tmax *= tmax_scale
    t = time.perf_counter()
    b = Batch(specs, device_ic=True)
    t_ic = time.perf_counter() - t
    st = b.run()
    t_run = st["kernel_ms"] / 1e3
    res = b.download(derivatives=False)
    b.close()
    t = time.perf_counter()
    Wk.initial_states(specs[:1024])
    t_host = (time.perf_counter() - t) * len(specs) / 1024
    out = {"workload": "C3: %d perturbations of the 5-body base system (P = 5, 8.1, 13.2, 21.6, 35.0 d), h = 0.1 d, "
                       "tmax 3000 d x%g, Jacobian + transit times + dt/delements" % (len(specs), tmax_scale),
           "system_steps": st["system_steps"], "transits": st["transits"], "device_s": t_run,
           "system_steps_per_s": st["system_steps"] / t_run, "transits_per_s": st["transits"] / t_run,
           "ic_on_gpu_s": t_ic, "ic_host_s_estimate": t_host,
           "ic_host_basis": "init_nbody in host processes for 1024 systems, scaled to the batch",
           "max_count": int(max(int(r["count"].max()) for r in res))}
    log("fit batch: %d systems, %d system-steps in %.2f s (%.3g/s), %d transits; IC on the GPU %.2f s "
        "(host estimate %.1f s)" % (len(specs), st["system_steps"], t_run, out["system_steps_per_s"], st["transits"],
                                    t_ic, t_host))
    return out


def cpu_baseline(specs, states, sample, threads, cap_steps=2000):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import oracle_lib as O
    from nbodygradient_amd.workloads import SystemSpec
    idx = list(range(min(sample, len(specs))))
    # cap each sampled system at `cap_steps` steps: the rate per system-step
    # does not depend on the run length, and the cap keeps threads balanced
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--tmax-scale", type=float, default=1.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--serve", type=int, default=1000, metavar="REQUESTS",
                    help="serving leg: the first REQUESTS requests of the C4 trace (0 = skip)")
    ap.add_argument("--serve-tmax-scale", type=float, default=0.1)
    ap.add_argument("--serve-chunk", type=int, default=128)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (read MIB, write MIB) through an elementwise "
                         "kernel: the reference reading for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--fit", type=int, default=16384, metavar="SYSTEMS",
                    help="C3 fit-style batch leg with initial conditions on the GPU (0: skip)")
    ap.add_argument("--fit-tmax-scale", type=float, default=0.1)
    ap.add_argument("--autotune", type=float, default=0.05, metavar="FRAC",
                    help="before the timed steps, pick the launch configuration with the autotuner on a copy of "
                         "the batch with step counts x FRAC (0: off; only with --coop auto)")
    ap.add_argument("--coop", default="auto", metavar="auto|off|K",
                    help="two-wave workgroups for the longest systems: automatic selection (default), off, or the "
                         "K longest systems")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (systems are independent: a "
              "multi-GPU run would partition the batch across GPUs by estimated cost)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import nbodygradient_amd as nb
    from nbodygradient_amd import workloads as Wk
    from nbodygradient_amd.batch import Batch

    t_setup = time.perf_counter()
    specs = Wk.c2(args.systems)
    if args.tmax_scale != 1.0:
        for s in specs:
            s.tmax *= args.tmax_scale
    states = Wk.initial_states(specs)
    ctx = nb.default_context()
    if args.coop == "off":
        ctx.set_coop(0, 0)
    elif args.coop != "auto":
        ctx.set_coop(2, int(args.coop))
    name, cus = ctx.info()
    log("device %s, %d CUs; %d systems prepared in %.1fs" % (name, cus, len(specs), time.perf_counter() - t_setup))

    # prime code loading with a tiny batch (not part of the workload)
    pb = Batch(specs[:8], states[:8])
    pb.run()
    pb.close()

    batch = Batch(specs, states, grad=True)
    tune = None
    if args.autotune > 0 and args.coop == "auto":
        t = time.perf_counter()
        tune = batch.autotune(args.autotune)
        c = tune["candidates"][tune["best"]]
        log("autotune (x%g copy, %.1f s): %s -> coop mode %d max %d, groups/CU %d" %
            (args.autotune, time.perf_counter() - t, ", ".join("%.0f" % q["ms"] for q in tune["candidates"]),
             c["coop_mode"], c["coop_max"], c["groups_per_cu"]))
    for w in range(args.warmup):
        batch.reset()
        batch.run()
    ms = []
    stats = None
    for k in range(args.steps):
        batch.reset()
        t = time.perf_counter()
        stats = batch.run()
        wall = time.perf_counter() - t
        ms.append(wall * 1e3)
        log("step %d: %.1f ms wall, %.1f ms device, %d system-steps, %d transits" %
            (k, wall * 1e3, stats["kernel_ms"], stats["system_steps"], stats["transits"]))
    ms_per_step = float(np.mean(ms))
    sys_steps = stats["system_steps"]
    value = sys_steps / (ms_per_step / 1e3)
    res = batch.download(derivatives=False)
    # per-launch roofline: algorithmic FP64 flops W(n) x units of the launch for
    # n (main steps + one per recorded transit) over that launch's duration,
    # timed with HIP events on the stream it ran on
    launches = {}
    for cls, L in stats["launches"].items():
        flops = sum(W(n) * u for n, u in L["units"].items())
        name = "nbg_run_coop<10>" if cls == 1 else "nbg_run<%d>" % cls  # class 1: two-wave workgroups
        launches[str(cls)] = {"kernel": name, "ms": L["ms"], "grid": L["grid"],
                              "units": sum(L["units"].values()),
                              "units_by_n": {str(n): u for n, u in sorted(L["units"].items())}, "flops": flops,
                              "tflops": flops / (L["ms"] / 1e3) / 1e12 if L["ms"] > 0 else 0.0}
    # the launches of a batch run concurrently and end together (each takes
    # its own class first, then smaller ones); the dominant one is the launch
    # that carries the most algorithmic work
    dom = max(launches, key=lambda k: launches[k]["flops"])
    achieved = launches[dom]["tflops"]
    total_flops = sum(L["flops"] for L in launches.values())
    log("download: %d systems, max count %d" % (len(res), max(int(r["count"].max()) for r in res)))
    for n in sorted(launches, key=int):
        L = launches[n]
        log("launch %s: grid %d, %.1f ms, %d units %s, %.3f TFLOP/s" % (L["kernel"], L["grid"], L["ms"], L["units"],
                                                                        L["units_by_n"], L["tflops"]))
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
        "config": {"workload": "C2: %d systems, n~U{3..9}, P log-U[2,200] d, h=P_in/30, tmax~U[500,4000] d%s, "
                               "Jacobian + transit times + dt/delements" %
                               (len(specs), "" if args.tmax_scale == 1.0 else " x%g" % args.tmax_scale),
                   "systems": len(specs), "system_steps": sys_steps, "transits": stats["transits"],
                   "newton_iters": stats["newton_iters"], "parallelism": "1 GPU, one persistent launch per body-count class (n <= 9 in one)",
                   "autotune": tune},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": HBM_BYTES_PER_UNIT * launches[dom]["units"],
                     "traffic_basis": "rocprofv3 FETCH_SIZE/WRITE_SIZE of nbg_run<9> on C2 x0.05, calibrated on nbg_copy_kernel: "
                                      "%.0f B per system-step x this launch's units" % HBM_BYTES_PER_UNIT,
                     "kernel": launches[dom]["kernel"], "launch_ms": launches[dom]["ms"],
                     "note": "FP64 compute bound (VALU); MI355X FP64 vector and matrix peaks are both 78.6 "
                             "TFLOP/s; W(n) = 2016 n^2 (n-1) + 126 n^3 flops per system-step"},
        "launches": launches,
        "batch_tflops": total_flops / (stats["kernel_ms"] / 1e3) / 1e12,
        "transits_per_s": stats["transits"] / (ms_per_step / 1e3),
    }
    if args.serve > 0:
        from nbodygradient_amd.server import Server, serve
        ctx.set_coop(1, 128)  # serving rounds keep the defaults, whatever the batch's autotuner chose
        ctx.set_launch(0)
        trace = Wk.c4(count=args.serve)
        arrivals = np.array([t for t, _ in trace])
        sspecs = [sp for _, sp in trace]
        for sp in sspecs:
            sp.tmax *= args.serve_tmax_scale
        sstates = Wk.initial_states(sspecs)
        sv = Server(max_bytes=32 << 30, chunk=args.serve_chunk)
        sres = serve(sv, sspecs, sstates, arrivals, log=log)
        sv.close()
        lat = sres["latency_s"]
        ssteps = sum(sp.nsteps for sp in sspecs)
        out["serving"] = {
            "workload": "C4 trace: first %d requests, Poisson arrivals at 2000/s, n~U{2..10}, tmax log-U[100,8000] d "
                        "x%g, h=P_in/30, Jacobian + transit times + dt/delements" % (len(sspecs), args.serve_tmax_scale),
            "chunk_steps": args.serve_chunk, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "max_s": float(np.max(lat)),
            "sustained_requests_per_s": len(sspecs) / sres["wall_s"], "wall_s": sres["wall_s"],
            "main_steps_per_s": ssteps / sres["wall_s"], "rounds": sres["rounds"],
            "offered_requests_per_s": len(sspecs) / float(arrivals[-1])}
        log("serving: %d requests, p50 %.3f s, p99 %.3f s, %.1f requests/s sustained, %.3g main steps/s" %
            (len(sspecs), out["serving"]["p50_s"], out["serving"]["p99_s"], out["serving"]["sustained_requests_per_s"],
             out["serving"]["main_steps_per_s"]))
    if args.fit > 0:
        out["fit_batch"] = fit_batch(args.fit, args.fit_tmax_scale)
    if not args.no_plain:
        lat = plain_latency(specs, states)
        t_plain = 0.0
        for sp, r in zip(specs, res):
            tr = int(np.minimum(r["count"], sp.ntt()).sum())
            t_plain += lat[sp.n] * (sp.nsteps + tr)
        plain_rate = sys_steps / t_plain
        out["plain_gpu"] = {"value": plain_rate, "unit": "system-steps/s",
                            "method": "one system per launch (nbg_integrate_transits); per-n step latency from 2 "
                                      "systems x <=1500 steps, summed over the batch"}
        out["batched_vs_plain"] = value / plain_rate
    if not args.no_cpu:
        # one thread per host core of this process's CPU share (at most 16)
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        done, secs, ns = cpu_baseline(specs, states, args.cpu_sample, threads)
        cpu_rate = done / secs
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "system-steps/s", "cores": threads, "kind": "port",
                               "sample": "first %d of the %d systems (seeded), each capped at 2000 steps, %d system-steps "
                                         "in %.1f s, OpenMP one system per thread" % (ns, len(specs), done, secs)}
        out["batched_vs_cpu"] = value / cpu_rate
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
