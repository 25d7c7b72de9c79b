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
        "cellsort": 4 * 6 + 32 + 24 + 8 + 4,
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


def timed_steps(eng, steps: int) -> float:
    eng.fetch()                            # drains the stream: barrier before timing
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.utotal_async()
    eng.fetch()                            # synchronises the stream (energies to host)
    return time.perf_counter() - t0


def run_ensemble(args):
    """Batched ensemble: 512 independent electrolyte / polyelectrolyte systems (N log-uniform
    in 200..20000, volume fraction 0.01, SPME 16^3..64^3 by box length, +-1/+-2 ions), all
    UTotal passes in one captured graph per step, pair tiles from the autotuner; 1000 MC
    trials per system batched across systems.  Beside it: the plain FP64 GPU path run one
    system at a time (eager kernels, a host synchronisation per system)."""
    from molsim_amd import autotune, systems
    t0 = time.perf_counter()
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
    tune = autotune.tune(engs, reps=3)
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

    # plain FP64 GPU path, one system at a time; its results are the reference for the
    # mixed batch's errors
    t_sum, t_de, n_de = 0.0, 0.0, 0
    ferr, eerr, deerr = [], [], []
    for s, sp in enumerate(specs):
        e64 = Engine(sp, precision="fp64", table=table)
        e64.utotal_async()
        e64.fetch()
        t0 = time.perf_counter()
        for _ in range(3):
            e64.utotal_async()
            e64.fetch()
        t_sum += (time.perf_counter() - t0) / 3
        r64 = e64.utotal()
        f64 = r64["force"]
        ferr.append(float(np.linalg.norm(res[s]["force"] - f64) / np.linalg.norm(f64)))
        eerr.append(abs(res[s]["u_tot"] - r64["u_tot"]) / (abs(r64["u_tot"]) + abs(r64["u_rec"])))
        k = args.fp64_trials
        t0 = time.perf_counter()
        for t in range(k):
            d = e64.dutotal(int(ips[s][t]), rns[s][t])
            if not d["overlap"]:
                deerr.append(abs(d["du_tot"] - de[s]["du_tot"][t]))
        t_de += time.perf_counter() - t0
        n_de += k
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
        "autotune": {str(b): {"n": v["n"], "tile": v["tile"],
                              "pair_us": {str(t): round(1e3 * ms, 2) for t, ms in v["ms"].items()}}
                     for b, v in sorted(tune.items())},
        "fp64_gpu_path": {"value": value64, "ms_per_ensemble_pass": 1e3 * t_sum,
                          "speedup_of_batched": value / value64,
                          "what": "plain FP64 path, eager, one system at a time"},
        "dE": {"trials_per_s": len(specs) * args.ntrial / dt_de,
               "fp64_one_at_a_time_trials_per_s": n_de / t_de,
               "speedup": (len(specs) * args.ntrial / dt_de) / (n_de / t_de),
               "max_abs_err_vs_fp64_kJmol": float(max(deerr)) if deerr else None},
        "err_vs_fp64_path": {"force_rel_l2_max": float(max(ferr)), "u_tot_rel_max": float(max(eerr))},
        "setup_s": {"generate": gen_s, "autotune": tune_s, "capture": cap_s},
    }


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
                    choices=["polyelectrolyte", "ensemble"])
    ap.add_argument("--nsys", type=int, default=512, help="ensemble: number of systems")
    ap.add_argument("--ntrial", type=int, default=1000, help="ensemble: MC trials per system")
    ap.add_argument("--fp64-trials", type=int, default=20,
                    help="ensemble: trials per system timed on the plain FP64 path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "ensemble":
        out = run_ensemble(args)
        if out:
            print(json.dumps(out), flush=True)
        return

    from molsim_amd import systems
    from molsim_amd.engine import Engine, build_table

    t_gen = time.perf_counter()
    sp = systems.polyelectrolyte()
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
                     "traffic": None, "alg_bytes_per_launch": nbytes},
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
                                         f"({tcpu:.1f} s, host has {os.cpu_count()} cpus)"}
        fr = ref["force"]
        out["err_vs_fp64_oracle"] = {
            "force_rel_l2": float(np.linalg.norm(res["force"] - fr) / np.linalg.norm(fr)),
            "force_rel_l2_fp64_path": float(np.linalg.norm(res64["force"] - fr) / np.linalg.norm(fr)),
            "u_tot_rel": abs(res["u_tot"] - ref["u_tot"]) / abs(ref["u_tot"]),
            "u_rec_rel": abs(res["u_rec"] - ref["u_rec"]) / abs(ref["u_rec"]),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
