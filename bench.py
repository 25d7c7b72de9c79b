"""SPME energy+force throughput of the MI355X non-bonded path on one GPU.

Workload (default): the polyelectrolyte solution -- 64 chains x 256 beads (q = -1) plus
16384 counterions, cubic box 400 A, SPME mesh 128^3 with order-6 B-splines, alpha 0.15/A,
rc 20 A cell list, tabulated LJ + Coulomb pair potential; mixed precision (FP32 pair,
spread and FFT arithmetic, FP64 accumulation).  One step = one full UTotal (cell list
build, pair loop, SPME spread / FFT / influence function / FFT / gather, reductions) with
the positions already resident in HBM.  value = particle-steps/s = N * steps / time.

Printed beside it (one JSON line): the plain FP64 single-system GPU path (the same kernels in
double, no graph), the force error of the timed path against the FP64 oracle, p50/p99
latency of synchronous UTotal requests, the roofline of the dominant kernel (event-timed on
the library's stream) and the oracle timed on the host CPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(name: str, n: int, K: int, fp64: bool) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of each kernel.

    Mesh kernels: FP32 real mesh 4 B/point, half-complex 8 B * H*K^2 (H = K/2+1), influence
    function 4 B * H^3.  Particle kernels: 24 B positions, 8 B charge, 4 B ids per particle
    read once; forces 24 B written (read-modify-write 48 B in the gather)."""
    r = 8 if fp64 else 4
    H = K // 2 + 1
    mesh_r = r * K ** 3
    mesh_c = 2 * r * H * K * K
    per_particle = {
        "bin": 24 + 8 + 4 + 4 + 3 * r + 4,
        "scatter": 4 + 4 + 4 + 4,
        "finalize": 4 + 24 + 4 + 32 + 4 + 4 + 4 + 4 * r + 4 + 4 + 4,
        "pair": 32 + 4 + 4 + 4 + 24,
        "spread": 4 + 4 + 3 * r + 8,
        "gather": 4 + 4 + 3 * r + 8 + 48,
    }
    mesh = {
        "spread": mesh_r,
        "fft_x": mesh_r + mesh_c,
        "fft_y": 2 * mesh_c,
        "fft_zconv": 2 * mesh_c + r * H ** 3,
        "fft_yi": 2 * mesh_c,
        "fft_xi": mesh_c + mesh_r,
        "gather": mesh_r,
    }
    return per_particle.get(name, 0) * n + mesh.get(name, 0)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_polyelectrolyte_32768_spme128.json")


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel of the default workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size), or None."""
    import re
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        data = json.load(f)
    for name, v in data.get("kernels", {}).items():
        if re.match(rf"k_{kernel}(_nb)?<", name) or name == f"k_{kernel}":
            return v["traffic_bytes"]
    return None


def timed_steps(eng, steps: int) -> float:
    eng.fetch()                            # drains the stream: barrier before timing
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.utotal_async()
    eng.fetch()                            # synchronises the stream (energies to host)
    return time.perf_counter() - t0


def run_ensemble(args, specs=None):
    """Batched ensemble: 512 independent electrolyte / polyelectrolyte systems (N log-uniform
    in 200..20000, volume fraction 0.01, SPME 16^3..64^3 by box length, +-1/+-2 ions), all
    UTotal passes in one captured graph per step, pair tiles from the autotuner; 1000 MC
    trials per system batched across systems.  Beside it: the plain FP64 GPU path run one
    system at a time (eager kernels, a host synchronisation per system)."""
    from molsim_amd import autotune, systems
    t0 = time.perf_counter()
    if specs is None:
        specs = systems.ensemble(args.nsys, workers=8)   # forked workers: before any GPU call
    gen_s = time.perf_counter() - t0
    from molsim_amd.engine import Batch, Engine, build_table
    table = build_table(specs[0].pot)
    ntot = sum(sp.n for sp in specs)
    print(f"# ensemble: {len(specs)} systems, {ntot} particles (generated in {gen_s:.1f} s)",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    engs = [Engine(sp, precision="mixed", table=table) for sp in specs]
    print(f"# {len(engs)} contexts in {time.perf_counter() - t0:.1f} s, "
          f"{sum(e.device_bytes() for e in engs) / 2**30:.2f} GiB resident",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tune = autotune.tune_batched(engs, reps=3)
    autotune.apply(engs, tune)
    tune_s = time.perf_counter() - t0
    print(f"# autotune {tune_s:.1f} s: " + ", ".join(
        f"b{b}(n~{v['n']})->{v['tile']}" for b, v in sorted(tune.items())), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    batch = Batch(engs)
    cap_s = time.perf_counter() - t0
    print(f"# batch graph captured in {cap_s:.2f} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        batch.utotal_async()
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.utotal_async()
    batch.sync()
    dt = time.perf_counter() - t0
    value = ntot * args.steps / dt
    print(f"# batched: {1e3 * dt / args.steps:.3f} ms per ensemble step, {value:.3e} particle-steps/s",
          file=sys.stderr, flush=True)
    if args.profile_only:
        return
    res = batch.utotal()

    # 1000 MC trials per system, one batched call
    rng = np.random.default_rng(7)
    ips, rns = [], []
    for sp in specs:
        ip = rng.integers(0, sp.n, size=args.ntrial)
        rn = sp.ro[ip] + rng.uniform(-1.0, 1.0, size=(args.ntrial, 3))
        L = sp.box[0]
        ips.append(ip)
        rns.append((rn + L / 2) % L - L / 2)
    batch.dutotal(ips, rns)
    t0 = time.perf_counter()
    de = batch.dutotal(ips, rns)
    dt_de = time.perf_counter() - t0
    print(f"# batched dE: {len(specs) * args.ntrial / dt_de:.3e} trials/s", file=sys.stderr, flush=True)

    # an MCPass per system, the chains side by side (after the trials above: it moves particles)
    from molsim_amd.engine import Random
    beta = systems.beta_kjmol()
    batch.mc_pass(8, 1.0, beta, [Random(-7 - s) for s in range(len(specs))], events=False)
    rngs = [Random(-1000 - s) for s in range(len(specs))]
    t0 = time.perf_counter()
    _, mc_acc, _ = batch.mc_pass(args.mc_pass_trials, 1.0, beta, rngs, events=False)
    dt_mc = time.perf_counter() - t0
    mc_rate = len(specs) * args.mc_pass_trials / dt_mc
    print(f"# batched MCPass: {mc_rate:.3e} trials/s ({int(mc_acc.sum())} accepted)", file=sys.stderr, flush=True)

    # plain FP64 GPU path, one system at a time; its results are the reference for the
    # mixed batch's errors
    t_sum, t_de, n_de = 0.0, 0.0, 0
    t_mc1, n_mc1 = 0.0, 0
    ferr, eerr, deerr = [], [], []
    for s, sp in enumerate(specs):
        e64 = Engine(sp, precision="fp64", table=table)
        e64.utotal_async()
        e64.fetch()
        # warm, best of 3 (a pass of the loop is short enough for host jitter to show)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            e64.utotal_async()
            e64.fetch()
            best = min(best, time.perf_counter() - t0)
        t_sum += best
        r64 = e64.utotal()
        f64 = r64["force"]
        ferr.append(float(np.linalg.norm(res[s]["force"] - f64) / np.linalg.norm(f64)))
        escale = abs(r64["u_tot"]) + abs(r64["u_rec"]) + float(np.abs(r64["u_twob"]).max())
        eerr.append(abs(res[s]["u_tot"] - r64["u_tot"]) / escale)
        k = args.fp64_trials
        t0 = time.perf_counter()
        for t in range(k):
            d = e64.dutotal(int(ips[s][t]), rns[s][t])
            if not d["overlap"]:
                deerr.append(abs(d["du_tot"] - de[s]["du_tot"][t]))
        t_de += time.perf_counter() - t0
        n_de += k
        # the same system's chain alone (FP64 device chain, one system at a time)
        t0 = time.perf_counter()
        e64.mc_pass(args.fp64_mc_trials, 1.0, beta, Random(-1000 - s))
        t_mc1 += time.perf_counter() - t0
        n_mc1 += args.fp64_mc_trials
        e64.close()
    value64 = ntot / t_sum
    print(f"# plain FP64 one at a time: {1e3 * t_sum:.2f} ms per ensemble pass, {value64:.3e} "
          f"particle-steps/s; speedup {value / value64:.1f}x", file=sys.stderr, flush=True)
    batch.close()
    for e in engs:
        e.close()
    nb = np.array([sp.n for sp in specs])
    return {
        "metric": "ensemble SPME E+F particle-steps/s (512 systems, one graph launch per step)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64acc",
        "data": "synthetic (seeded electrolytes and random-walk polyelectrolytes)",
        "config": {"workload": f"ensemble_{len(specs)}", "systems": len(specs), "particles": int(ntot),
                   "n_min": int(nb.min()), "n_max": int(nb.max()),
                   "meshes": {str(k): int((np.array([sp.nmesh for sp in specs]) == k).sum())
                              for k in (16, 32, 64)},
                   "rcut": specs[0].rcut, "alpha": specs[0].alpha, "order": specs[0].order,
                   "parallelism": "single GPU"},
        "autotune": {str(b): {"n_median": v["n"], "systems": v["members"], "tile": v["tile"],
                              "bucket_pass_ms": {str(t): round(ms, 3) for t, ms in v["ms"].items()}}
                     for b, v in sorted(tune.items())},
        "fp64_gpu_path": {"value": value64, "ms_per_ensemble_pass": 1e3 * t_sum,
                          "speedup_of_batched": value / value64,
                          "what": "plain FP64 path, eager, one system at a time, warm, best of 3 "
                                  "per system"},
        "dE": {"trials_per_s": len(specs) * args.ntrial / dt_de,
               "fp64_one_at_a_time_trials_per_s": n_de / t_de,
               "speedup": (len(specs) * args.ntrial / dt_de) / (n_de / t_de),
               "max_abs_err_vs_fp64_kJmol": float(max(deerr)) if deerr else None},
        "mc_pass": {"trials_per_s": mc_rate, "trials_per_system": args.mc_pass_trials,
                    "accepted": int(mc_acc.sum()), "dtran": 1.0,
                    "fp64_one_system_at_a_time_trials_per_s": n_mc1 / t_mc1,
                    "speedup": mc_rate / (n_mc1 / t_mc1),
                    "what": "one device chain per system side by side (msg_batch_mc_pass), against "
                            "each system's FP64 device chain run alone, one system after another"},
        "err_vs_fp64_path": {"force_rel_l2_max": float(max(ferr)), "u_tot_err_max": float(max(eerr)),
                             "u_tot_err_scale": "|u_tot| + |u_rec| + max |u_twob| of the FP64 result"},
        "setup_s": {"generate": gen_s, "autotune": tune_s, "capture": cap_s},
    }


def run_serve(args, pool=None):
    """Served request stream: a pool of resident systems (the polyelectrolyte composition
    rescaled to N log-uniform in 100..100000), Poisson arrivals of requests for energy +
    forces or for 256 single-move trial energies, continuous batching over two execution
    units with HBM-footprint admission.  Beside it: the plain FP64 one-at-a-time path (each
    request alone, trials one DUTotal call each), served end to end in real time at the loads
    below its capacity, and as the FIFO recursion over its measured service times at every
    load (above capacity its queue grows without bound)."""
    from molsim_amd import serve, systems
    t0 = time.perf_counter()
    if pool is None:
        pool = systems.served_pool(args.nsys_serve, workers=8)     # forked: before any GPU call
    gen_s = time.perf_counter() - t0
    from molsim_amd.engine import Engine, build_table
    table = build_table(pool[0].pot)
    ns = np.array([sp.n for sp in pool.values()])
    print(f"# serve: {len(pool)} systems, N {ns.min()}..{ns.max()} (generated in {gen_s:.1f} s)",
          file=sys.stderr, flush=True)

    # plain FP64 path: service time of each (system, kind), median of 3; the contexts stay
    # resident for the end-to-end run below
    rng = np.random.default_rng(11)
    svc = {}
    e64s = {}
    for sid, sp in pool.items():
        e = Engine(sp, precision="fp64", table=table)
        e64s[sid] = e
        e.utotal()
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            e.utotal()
            ts.append(time.perf_counter() - t1)
        svc[(sid, "ef")] = float(np.median(ts))
        ips = rng.integers(0, sp.n, size=args.ntrial_serve)
        rn = sp.ro[ips] + rng.uniform(-1.0, 1.0, size=(args.ntrial_serve, 3))
        e.dutotal(int(ips[0]), rn[0])          # one-time trial set-up is not request time
        ts = []
        for _ in range(2):
            t1 = time.perf_counter()
            for k in range(len(ips)):
                e.dutotal(int(ips[k]), rn[k])
            ts.append(time.perf_counter() - t1)
        svc[(sid, "de")] = float(np.median(ts))
    mean_svc = 0.5 * np.mean([v for (s_, k), v in svc.items() if k == "ef"]) + \
        0.5 * np.mean([v for (s_, k), v in svc.items() if k == "de"])
    cap64 = 1.0 / mean_svc
    print(f"# plain FP64 path: mean service {1e3 * mean_svc:.2f} ms -> capacity {cap64:.1f} req/s",
          file=sys.stderr, flush=True)

    def fifo(reqs):
        end, lat = 0.0, []
        for r in sorted(reqs, key=lambda q: q.t_arrival):
            start = max(r.t_arrival, end)
            end = start + svc[(r.sys_id, r.kind)]
            lat.append(end - r.t_arrival)
        lat = np.array(lat) * 1e3
        return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                "throughput_rps": len(reqs) / max(end, reqs[-1].t_arrival)}

    # the same path served end to end in real time at the loads it can sustain (below its
    # capacity; above it the queue grows without bound and only the recursion is reported)
    e2e = {}
    for mult in args.qps_mult:
        if mult >= 1.0:
            continue
        reqs = serve.request_stream(pool, qps=mult * cap64, duration=args.duration, seed=int(100 * mult),
                                    frac_ef=0.5, ntrial=args.ntrial_serve)
        st = serve.serve_serial(e64s, reqs)
        e2e[mult] = st
        print(f"# plain FP64 path served end to end at {mult:.2f}x capacity: p50 {st['latency_ms']['p50']:.2f} "
              f"p99 {st['latency_ms']['p99']:.2f} ms ({st['requests']} requests)", file=sys.stderr, flush=True)
    for e in e64s.values():
        e.close()
    e64s.clear()

    srv = serve.Server(pool, table, hbm_budget=int(args.hbm_gb * 2 ** 30), units=args.units,
                       max_per_unit=64)
    # residency warm-up (not timed): one energy request and one trial request per system
    warm = serve.request_stream(pool, qps=1e6, duration=1e-3, seed=1, ntrial=4)[:0]
    for i, (sid, sp) in enumerate(pool.items()):
        warm.append(serve.Request(-1 - 2 * i, sid, "ef", 0.0))
        warm.append(serve.Request(-2 - 2 * i, sid, "de", 0.0, ips=np.arange(4) % sp.n,
                                  rnew=sp.ro[np.arange(4) % sp.n].copy()))
    srv.run(warm)
    srv.done.clear()
    print(f"# resident: {len(srv.resident)} systems, {srv.bytes / 2 ** 30:.2f} GiB admitted",
          file=sys.stderr, flush=True)
    levels = []
    for mult in args.qps_mult:
        qps = mult * cap64
        reqs = serve.request_stream(pool, qps=qps, duration=args.duration, seed=int(100 * mult),
                                    frac_ef=0.5, ntrial=args.ntrial_serve)
        base = fifo([serve.Request(r.rid, r.sys_id, r.kind, r.t_arrival) for r in reqs])
        st = srv.run(reqs)
        srv.done.clear()
        lv = {"qps": qps, "qps_over_fp64_capacity": mult, "requests": st["requests"],
              "served": {"p50_ms": st["latency_ms"]["p50"], "p99_ms": st["latency_ms"]["p99"],
                         "throughput_rps": st["throughput_rps"],
                         "p99_ms_ef": st.get("latency_ms_ef", {}).get("p99"),
                         "p99_ms_de": st.get("latency_ms_de", {}).get("p99")},
              "fp64_one_at_a_time": {"p50_ms": base["p50"], "p99_ms": base["p99"],
                                     "throughput_rps": base["throughput_rps"],
                                     "stable": mult < 1.0,
                                     "what": "FIFO recursion over the measured service times"}}
        if mult in e2e:
            st64 = e2e[mult]
            lv["fp64_one_at_a_time_end_to_end"] = {
                "p50_ms": st64["latency_ms"]["p50"], "p99_ms": st64["latency_ms"]["p99"],
                "throughput_rps": st64["throughput_rps"], "requests": st64["requests"],
                "what": "served in real time, one request at a time on resident FP64 contexts"}
        levels.append(lv)
        print(f"# qps {qps:9.1f}: served p50 {st['latency_ms']['p50']:.3f} p99 {st['latency_ms']['p99']:.3f} ms "
              f"({st['throughput_rps']:.0f} rps) | fp64 serial p50 {base['p50']:.2f} p99 {base['p99']:.2f} ms",
              file=sys.stderr, flush=True)
    evictions, deferred = srv.evictions, srv.deferred
    srv.close()
    main_lv = levels[min(1, len(levels) - 1)]
    return {
        "metric": "served SPME E+F / 256-move dE requests: p50/p99 latency and throughput at fixed QPS",
        "value": main_lv["served"]["throughput_rps"], "unit": "requests/s", "n_gpus": 1,
        "steps": main_lv["requests"], "warmup": len(pool), "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64acc",
        "data": "synthetic (polyelectrolyte composition rescaled, seeded Poisson arrivals)",
        "config": {"workload": f"served_{len(pool)}", "systems": len(pool), "n_min": int(ns.min()),
                   "n_max": int(ns.max()), "ntrial": args.ntrial_serve, "frac_ef": 0.5,
                   "units": args.units, "hbm_budget_gib": args.hbm_gb, "duration_s": args.duration,
                   "parallelism": "single GPU"},
        "levels": levels,
        "fp64_capacity_rps": cap64,
        "evictions": evictions, "deferred": deferred,
    }


def run_electrolyte_mc(args):
    """The 1000-ion primitive-model electrolyte (std Ewald, rc = L/2, all pairs, FP64):
    UTotal with forces, then 10^4 single-particle MC trials of at most 1 A per axis, as one
    MCPass on the device (SPartMove / Metropolis / MCUpdate with MOLSIM's generator), and the
    same 10^4 trials as independent DUTotal evaluations against the start configuration.
    CPU leg: the oracle's DUTotal on a bounded sample of the trials."""
    from molsim_amd import systems
    from molsim_amd.engine import Engine, Random, build_table
    import oracle
    sp = systems.electrolyte_primitive()
    table = build_table(sp.pot)
    beta = systems.beta_kjmol()
    out = {}
    for prec in ("fp64", "mixed"):
        eng = Engine(sp, precision=prec, table=table)
        res = eng.utotal()
        t0 = time.perf_counter()
        nrep = 20
        for _ in range(nrep):
            eng.utotal()
        t_ut = (time.perf_counter() - t0) / nrep
        rng = np.random.default_rng(5)
        ips = rng.integers(0, sp.n, size=args.mc_trials)
        rn = sp.ro[ips] + rng.uniform(-1.0, 1.0, size=(args.mc_trials, 3))
        rn = (rn + sp.box[0] / 2) % sp.box[0] - sp.box[0] / 2
        eng.dutotal_batch(ips[:16], rn[:16])
        t0 = time.perf_counter()
        de = eng.dutotal_batch(ips, rn)
        t_de = time.perf_counter() - t0
        gen = Random(-12345)
        t0 = time.perf_counter()
        dsum, nacc, ev = eng.mc_pass(args.mc_trials, 2.0, beta, gen, isamp=0)
        t_mc = time.perf_counter() - t0
        out[prec] = {"utotal_ms": 1e3 * t_ut, "dE_batch_trials_per_s": args.mc_trials / t_de,
                     "mc_pass_trials_per_s": args.mc_trials / t_mc, "mc_accepted": int(nacc),
                     "u_tot": res["u_tot"], "force": res["force"], "de": de}
        print(f"# {prec}: UTotal {1e3 * t_ut:.3f} ms, batched dE {args.mc_trials / t_de:.3e} trials/s, "
              f"MCPass {args.mc_trials / t_mc:.3e} trials/s ({nacc} accepted)", file=sys.stderr, flush=True)
        eng.close()
    # oracle: forces, then the trials the way MOLSIM evaluates them (resident structure
    # factors updated per trial in O(nkvec), DUTwoBodyEwaldRecStd), time and error
    otab = oracle.build_table(sp.pot)
    osys = oracle.System(sp.ro, sp.ipt, sp.az, sp.box, sp.rcut, sp.pot["epsi4pi"], otab, npt=sp.npt,
                         list_mode=0, ewald=1, surf=sp.lsurf, alpha=sp.alpha, ncut=sp.ncut)
    ref = osys.utotal()
    rng = np.random.default_rng(5)
    ips = rng.integers(0, sp.n, size=args.mc_trials)
    rn = sp.ro[ips] + rng.uniform(-1.0, 1.0, size=(args.mc_trials, 3))
    rn = (rn + sp.box[0] / 2) % sp.box[0] - sp.box[0] / 2
    trials = osys.trials()
    t0 = time.perf_counter()
    nref, errs = 0, {"fp64": 0.0, "mixed": 0.0}
    while nref < args.mc_trials and time.perf_counter() - t0 < args.cpu_seconds:
        r = trials.dutotal(int(ips[nref]), rn[nref])
        if not r["overlap"]:
            for prec in errs:
                errs[prec] = max(errs[prec], abs(out[prec]["de"]["du_tot"][nref] - r["du_tot"]))
        nref += 1
    t_cpu = time.perf_counter() - t0
    trials.close()
    fr = ref["force"]
    spme = run_electrolyte_mc_spme(args, sp, table, beta)
    return {
        "metric": "primitive electrolyte: DUTotal trials/s (10^4 single-particle moves), std Ewald FP64",
        "value": out["fp64"]["dE_batch_trials_per_s"], "unit": "trials/s", "n_gpus": 1,
        "steps": args.mc_trials, "warmup": 16, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded uniform insertion with overlap rejection)",
        "config": {"workload": "electrolyte_1000_stdewald", "n": sp.n, "box": float(sp.box[0]),
                   "alpha": sp.alpha, "ncut": sp.ncut, "rcut": sp.rcut, "list": "all pairs",
                   "parallelism": "single GPU"},
        "paths": {p: {k: v for k, v in d.items() if k not in ("force", "de")} for p, d in out.items()},
        "err_vs_fp64_oracle": {
            "force_rel_l2_fp64": float(np.linalg.norm(out["fp64"]["force"] - fr) / np.linalg.norm(fr)),
            "force_rel_l2_mixed": float(np.linalg.norm(out["mixed"]["force"] - fr) / np.linalg.norm(fr)),
            "dE_max_abs_fp64_kJmol": errs["fp64"], "dE_max_abs_mixed_kJmol": errs["mixed"],
            "dE_trials_checked": nref},
        "cpu_baseline": {"value": nref / t_cpu, "unit": "trials/s", "cores": 1, "kind": "port",
                         "sample": f"{nref} oracle DUTotal trials ({t_cpu:.1f} s) with the structure "
                                   f"factors resident and updated per trial in O(nkvec) "
                                   f"(DUTwoBodyEwaldRecStd), all-pairs real space",
                         "cpu": cpu_model()},
        "spme_chain": spme,
    }


def run_electrolyte_mc_spme(args, sp, table, beta):
    """The same electrolyte with SPME on the reference decks' mesh (40^3, order 5): one MCPass
    of up to 10^4 trials on the device (the chain corrects phi for its accepted moves and
    recomputes it by a UTotal pass after every 16), FP64 and mixed.  CPU leg: the oracle's
    DUTotal, which spreads the moved charge and transforms the whole mesh per trial as
    DUTwoBodyEwaldRecSPM does, on a bounded sample; the GPU's batched trials are checked
    against it."""
    from molsim_amd import systems
    from molsim_amd.engine import Engine, Random
    import oracle
    sps = systems.SystemSpec(**{**sp.__dict__, "ewald": "spm", "nmesh": 40, "order": 5})
    ntrial = min(args.mc_trials, 10000)
    rng = np.random.default_rng(7)
    ips = rng.integers(0, sps.n, size=64)
    rn = sps.ro[ips] + rng.uniform(-1.0, 1.0, size=(64, 3))
    rn = (rn + sps.box[0] / 2) % sps.box[0] - sps.box[0] / 2
    paths, des = {}, {}
    for prec in ("fp64", "mixed"):
        eng = Engine(sps, precision=prec, table=table)
        eng.utotal()
        des[prec] = eng.dutotal_batch(ips, rn)["du_tot"]
        eng.mc_pass(64, 2.0, beta, Random(-999), isamp=0)  # warm: kernels loaded
        eng.utotal()
        gen = Random(-12345)
        t0 = time.perf_counter()
        _, nacc, _ = eng.mc_pass(ntrial, 2.0, beta, gen, isamp=0)
        t_mc = time.perf_counter() - t0
        paths[prec] = {"mc_pass_trials_per_s": ntrial / t_mc, "mc_accepted": int(nacc), "trials": ntrial}
        print(f"# SPME 40^3 p5 {prec}: MCPass {ntrial / t_mc:.3e} trials/s ({nacc} accepted)",
              file=sys.stderr, flush=True)
        eng.close()
    otab = oracle.build_table(sps.pot)
    osys = oracle.System(sps.ro, sps.ipt, sps.az, sps.box, sps.rcut, sps.pot["epsi4pi"], otab, npt=sps.npt,
                         list_mode=0, ewald=2, surf=sps.lsurf, alpha=sps.alpha, K=40, order=5)
    osys.utotal()
    t0 = time.perf_counter()
    nref, err = 0, {"fp64": 0.0, "mixed": 0.0}
    while nref < len(ips) and time.perf_counter() - t0 < args.cpu_seconds:
        r = osys.dutotal(int(ips[nref]), rn[nref])
        if not r["overlap"]:
            for prec in err:
                err[prec] = max(err[prec], abs(des[prec][nref] - r["du_tot"]))
        nref += 1
    t_cpu = time.perf_counter() - t0
    return {"config": {"workload": "electrolyte_1000_spme40", "n": sps.n, "box": float(sps.box[0]),
                       "alpha": sps.alpha, "nmesh": 40, "order": 5, "rcut": sps.rcut, "list": "all pairs"},
            "paths": paths,
            "dE_max_abs_vs_oracle_kJmol": {**err, "trials_checked": nref},
            "cpu_baseline": {"value": nref / t_cpu, "unit": "trials/s", "cores": 1, "kind": "port",
                             "sample": f"{nref} oracle DUTotal trials ({t_cpu:.1f} s), each spreading the "
                                       f"charges and transforming the 40^3 mesh (DUTwoBodyEwaldRecSPM)",
                             "fft": "the oracle's own mixed-radix complex 3-D FFT (not an FFTW-class "
                                    "transform; MOLSIM links FFTW)",
                             "cpu": cpu_model()}}


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), for the CPU baseline's record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_ionic_md(args):
    """The 1,048,576-particle ionic fluid: SPME 256^3 order 6, rc 18 A Verlet list with a
    2 A skin, velocity-Verlet MD on the device (dt 0.005, masses 20 g/mol, velocities at
    298 K), the list rebuilt on skin violation.  value = particle-steps/s over the timed
    steps.  Beside it: the plain FP64 path over a few steps, and one oracle UTotal of the
    same system on the host (CPU leg)."""
    from molsim_amd import systems
    from molsim_amd.engine import Engine, build_table
    t0 = time.perf_counter()
    sp = systems.ionic_fluid()
    gen_s = time.perf_counter() - t0
    print(f"# {sp.name}: N={sp.n} L={sp.box[0]:.1f} (generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    table = build_table(sp.pot)
    mass = [20.0, 20.0]
    rng = np.random.default_rng(4)
    kT = 2.479  # kJ/mol at 298 K
    v0 = rng.normal(size=sp.ro.shape) * np.sqrt(kT * 100.0 / 20.0)
    res = {}
    f0 = {}
    for prec, nsteps in (("mixed", args.md_steps), ("fp64", args.md_fp64_steps)):
        eng = Engine(sp, precision=prec, table=table, force_scale=100.0)
        f0[prec] = eng.utotal()["force"]          # start configuration, both paths
        eng.md_init(mass, v0)
        eng.md_run(args.md_warmup, 0.005, log=False)
        b0 = eng.nlist_builds()
        t0 = time.perf_counter()
        log = eng.md_run(nsteps, 0.005, log=True)
        dt = time.perf_counter() - t0
        etot = log[:, 0] * 100.0 + log[:, 1]
        res[prec] = {"particle_steps_per_s": sp.n * nsteps / dt, "ms_per_step": 1e3 * dt / nsteps,
                     "steps": nsteps, "list_builds": eng.nlist_builds() - b0,
                     "energy_drift_rel_K": float(np.abs(etot - etot[0]).max() / abs(log[0, 1]))}
        print(f"# {prec}: {1e3 * dt / nsteps:.3f} ms/step, {res[prec]['list_builds']} rebuilds",
              file=sys.stderr, flush=True)
        eng.close()
    ferr = float(np.linalg.norm(f0["mixed"] - f0["fp64"]) / np.linalg.norm(f0["fp64"]))
    out = {
        "metric": "ionic fluid MD: SPME E+F particle-steps/s (velocity Verlet, Verlet list)",
        "value": res["mixed"]["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.md_steps, "warmup": args.md_warmup, "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64acc",
        "data": "synthetic (seeded uniform insertion with overlap rejection, Maxwell velocities)",
        "config": {"workload": sp.name, "n": sp.n, "box": float(sp.box[0]), "nmesh": sp.nmesh,
                   "order": sp.order, "rcut": sp.rcut, "skin": sp.drnlist, "dt": 0.005,
                   "parallelism": "single GPU"},
        "paths": res,
        "err_vs_fp64_path": {"force_rel_l2": ferr, "what": "UTotal forces at the start configuration"},
    }
    if not args.no_cpu:
        import oracle
        otab = oracle.build_table(sp.pot)
        osys = oracle.System(sp.ro, sp.ipt, sp.az, sp.box, sp.rcut, sp.pot["epsi4pi"], otab,
                             npt=sp.npt, list_mode=1, ewald=2, surf=sp.lsurf, alpha=sp.alpha,
                             K=sp.nmesh, order=sp.order)
        if not getattr(args, "md_omp_only", False):
            t0 = time.perf_counter()
            osys.utotal()
            tcpu = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": sp.n / tcpu, "unit": "particle-steps/s", "cores": 1,
                                   "kind": "port", "sample": f"one oracle UTotal of the full system ({tcpu:.1f} s)",
                                   "cpu": cpu_model()}
        # the same on all host cores (the oracle's OpenMP build)
        t0 = time.perf_counter()
        osys.utotal(omp=True)
        tomp = time.perf_counter() - t0
        out["cpu_baseline_all_cores"] = {"value": sp.n / tomp, "unit": "particle-steps/s",
                                         "cores": oracle.threads(omp=True), "kind": "port",
                                         "sample": f"one UTotal of the full system by the oracle's OpenMP "
                                                   f"build ({tomp:.1f} s)", "cpu": cpu_model()}
    return out


def run_copy(args):
    """A plain device-to-device copy of a 1 GiB buffer (torch's copy kernel): the measured HBM
    bandwidth beside the 8 TB/s peak, and the reference for scaling the profiler's
    FETCH_SIZE / WRITE_SIZE counters (a copy reads and writes exactly its size)."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(args.steps, 1)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nbytes = 2 * 4 * n
    return {"metric": "device copy bandwidth (read + write)", "value": nbytes / dt / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": reps, "warmup": 3, "ms_per_step": 1e3 * dt, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "copy_1GiB", "bytes_read": 4 * n, "bytes_written": 4 * n},
            "peak_GBps": HBM_PEAK_GBS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps of the timed path only (for rocprofv3)")
    ap.add_argument("--workload", default="polyelectrolyte",
                    choices=["polyelectrolyte", "ensemble", "serve", "electrolyte_mc", "ionic_md", "copy"])
    ap.add_argument("--mc-trials", type=int, default=10000, help="electrolyte_mc: MC trials")
    ap.add_argument("--md-steps", type=int, default=1000, help="ionic_md: timed MD steps")
    ap.add_argument("--md-warmup", type=int, default=10, help="ionic_md: untimed MD steps")
    ap.add_argument("--md-fp64-steps", type=int, default=20, help="ionic_md: FP64 path steps")
    ap.add_argument("--nsys-serve", type=int, default=48, help="serve: resident pool size")
    ap.add_argument("--ntrial-serve", type=int, default=256, help="serve: moves per dE request")
    ap.add_argument("--units", type=int, default=2, help="serve: execution units in flight")
    ap.add_argument("--hbm-gb", type=float, default=8.0, help="serve: HBM budget (GiB)")
    ap.add_argument("--duration", type=float, default=3.0, help="serve: seconds per QPS level")
    ap.add_argument("--qps-mult", type=float, nargs="+", default=[0.6, 6.0, 30.0],
                    help="serve: QPS levels as multiples of the FP64 path's capacity")
    ap.add_argument("--nsys", type=int, default=512, help="ensemble: number of systems")
    ap.add_argument("--ens-steps", type=int, default=30, help="ensemble steps in the default run")
    ap.add_argument("--quick", action="store_true",
                    help="default workload only (no ensemble / served-stream summaries)")
    ap.add_argument("--ntrial", type=int, default=1000, help="ensemble: MC trials per system")
    ap.add_argument("--fp64-trials", type=int, default=20,
                    help="ensemble: trials per system timed on the plain FP64 path")
    ap.add_argument("--mc-pass-trials", type=int, default=200,
                    help="ensemble: MCPass trials per system (chains side by side)")
    ap.add_argument("--fp64-mc-trials", type=int, default=40,
                    help="ensemble: MCPass trials per system of the FP64 chain run alone")
    args = ap.parse_args()
    if args.profile_only:
        args.quick = True
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload != "polyelectrolyte":
        out = {"ensemble": run_ensemble, "serve": run_serve, "electrolyte_mc": run_electrolyte_mc,
               "ionic_md": run_ionic_md, "copy": run_copy}[args.workload](args)
        if out:
            print(json.dumps(out), flush=True)
        return

    from molsim_amd import systems
    from molsim_amd.engine import Engine, build_table

    t_gen = time.perf_counter()
    sp = systems.polyelectrolyte()
    # the ensemble and the served pool are generated here, in forked workers, before any
    # GPU call of this process
    ens_specs = None if args.quick else systems.ensemble(args.nsys, workers=8)
    serve_pool = None if args.quick else systems.served_pool(args.nsys_serve, workers=8)
    table = build_table(sp.pot)
    print(f"# workload {sp.name}: N={sp.n} K={sp.nmesh} order={sp.order} "
          f"(generated in {time.perf_counter() - t_gen:.1f} s)", file=sys.stderr, flush=True)

    eng = Engine(sp, precision="mixed", table=table)
    for _ in range(args.warmup):
        eng.utotal_async()
    eng.fetch()
    dt = timed_steps(eng, args.steps)
    value = sp.n * args.steps / dt
    ms_per_step = 1e3 * dt / args.steps
    print(f"# mixed: {ms_per_step:.4f} ms/step", file=sys.stderr, flush=True)
    if args.profile_only:
        eng.close()
        return

    # forces of the timed path for the error against the FP64 oracle
    res = eng.utotal()

    # per-kernel event timing (eager launches, same kernels) for the roofline
    eng.profile(True)
    nprof = max(20, args.steps // 5)
    for _ in range(nprof):
        eng.utotal_async()
    eng.fetch()
    prof = eng.profile_read()
    eng.profile(False)
    kernels = {k: {"avg_us": 1e3 * ms / n, "launches": n} for k, (ms, n) in prof.items()}
    dominant = max(kernels, key=lambda k: kernels[k]["avg_us"])
    nbytes = kernel_bytes(dominant, sp.n, sp.nmesh, False)
    achieved = nbytes / (kernels[dominant]["avg_us"] * 1e-6) / 1e9
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["avg_us"]):
        b = kernel_bytes(k, sp.n, sp.nmesh, False)
        v["alg_bytes"] = b
        v["alg_GBps"] = b / (v["avg_us"] * 1e-6) / 1e9 if b else None
        print(f"#   {k:10s} {v['avg_us']:9.2f} us  {b / 1e6:8.2f} MB  "
              f"{(v['alg_GBps'] or 0):8.1f} GB/s", file=sys.stderr, flush=True)

    # latency of synchronous requests (UTotal with forces returned to host memory)
    lat = []
    for _ in range(50):
        t0 = time.perf_counter()
        eng.utotal()
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3

    # plain FP64 single-system GPU path
    eng64 = Engine(sp, precision="fp64", table=table)
    for _ in range(5):
        eng64.utotal_async()
    eng64.fetch()
    dt64 = timed_steps(eng64, args.fp64_steps)
    value64 = sp.n * args.fp64_steps / dt64
    res64 = eng64.utotal()
    eng64.close()
    eng.close()

    out = {
        "metric": "SPME E+F particle-steps/s at <=1e-5 rel err vs FP64; p50/p99 request latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64acc",
        "data": "synthetic (seeded random-walk chains + counterions)",
        "config": {"workload": sp.name, "n": sp.n, "box": float(sp.box[0]), "nmesh": sp.nmesh,
                   "order": sp.order, "alpha": sp.alpha, "rcut": sp.rcut, "list": sp.list_mode,
                   "parallelism": "single GPU"},
        "fp64_gpu_path": {"value": value64, "ms_per_step": 1e3 * dt64 / args.fp64_steps,
                          "speedup_of_mixed": value / value64},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": "synchronous UTotal incl. force copy to host, single system"},
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in kernels.items()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dominant), "alg_bytes_per_launch": nbytes,
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)},
    }

    if not args.no_cpu:
        # cpu_baseline leg: the FP64 oracle (a single-threaded restatement of MOLSIM's path)
        # timed on a bounded sample of the same workload; its forces are also the FP64
        # reference for the error of the two GPU paths.
        import oracle
        otab = oracle.build_table(sp.pot)
        osys = oracle.System(sp.ro, sp.ipt, sp.az, sp.box, sp.rcut, sp.pot["epsi4pi"], otab,
                             npt=sp.npt, list_mode=1, ewald=2, surf=sp.lsurf, alpha=sp.alpha,
                             K=sp.nmesh, order=sp.order)
        t0 = time.perf_counter()
        nrun = 0
        while True:
            ref = osys.utotal()
            nrun += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        tcpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sp.n * nrun / tcpu, "unit": "particle-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{nrun} full UTotal evaluations of the same workload "
                                         f"({tcpu:.1f} s, host has {os.cpu_count()} cpus)",
                               "fft": "the oracle's own radix-2 complex 3-D FFT (not an FFTW-class "
                                      "transform; MOLSIM links FFTW)",
                               "cpu": cpu_model()}
        # the same on all host cores (the oracle's OpenMP build), a sample of half the time
        t0 = time.perf_counter()
        nomp = 0
        while True:
            osys.utotal(omp=True)
            nomp += 1
            if time.perf_counter() - t0 > args.cpu_seconds / 2:
                break
        tomp = time.perf_counter() - t0
        out["cpu_baseline_all_cores"] = {"value": sp.n * nomp / tomp, "unit": "particle-steps/s",
                                         "cores": oracle.threads(omp=True), "kind": "port",
                                         "sample": f"{nomp} full UTotal evaluations by the oracle's OpenMP "
                                                   f"build ({tomp:.1f} s)", "cpu": cpu_model()}
        fr = ref["force"]
        out["err_vs_fp64_oracle"] = {
            "force_rel_l2": float(np.linalg.norm(res["force"] - fr) / np.linalg.norm(fr)),
            "force_rel_l2_fp64_path": float(np.linalg.norm(res64["force"] - fr) / np.linalg.norm(fr)),
            "u_tot_rel": abs(res["u_tot"] - ref["u_tot"]) / abs(ref["u_tot"]),
            "u_rec_rel": abs(res["u_rec"] - ref["u_rec"]) / abs(ref["u_rec"]),
        }
    if not args.quick:
        # the batched ensemble and the served request stream, summarised in the same line
        ens = run_ensemble(args_for(args, steps=args.ens_steps, warmup=3), ens_specs)
        out["ensemble"] = {k: ens[k] for k in ("value", "unit", "ms_per_step", "fp64_gpu_path", "dE",
                                              "mc_pass", "err_vs_fp64_path", "config")}
        # the batched ensemble against the FP64 path's single-system rate on the headline
        # workload (the FP64 path at its best, not one small system at a time)
        out["ensemble"]["speedup_vs_fp64_single_system_rate"] = ens["value"] / value64
        srv = run_serve(args_for(args, qps_mult=args.qps_mult[:2], duration=2.0), serve_pool)
        out["serve"] = {k: srv[k] for k in ("levels", "fp64_capacity_rps", "config")}
        # the 1000-ion std-Ewald electrolyte (MC trials) and the 1,048,576-ion fluid (MD)
        mc = run_electrolyte_mc(args_for(args, cpu_seconds=min(args.cpu_seconds, 5.0)))
        out["electrolyte_mc"] = {k: mc[k] for k in ("value", "unit", "paths", "err_vs_fp64_oracle",
                                                    "cpu_baseline", "config", "spme_chain")}
        md = run_ionic_md(args_for(args, md_steps=100, md_warmup=5, md_fp64_steps=5, md_omp_only=True))
        out["ionic_md"] = {k: md[k] for k in ("value", "unit", "ms_per_step", "paths", "err_vs_fp64_path",
                                              "config") if k in md}
        if "cpu_baseline_all_cores" in md:
            out["ionic_md"]["cpu_baseline_all_cores"] = md["cpu_baseline_all_cores"]
    print(json.dumps(out), flush=True)


def args_for(args, **kw):
    a = argparse.Namespace(**vars(args))
    for k, v in kw.items():
        setattr(a, k, v)
    return a


if __name__ == "__main__":
    main()
