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
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

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
