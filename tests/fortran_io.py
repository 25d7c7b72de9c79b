"""Writes the input of molsim_amd/fortran/drop_in_driver and parses its output."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "molsim_amd", "fortran", "drop_in_driver")


def write_input(path, sp, trials):
    p = sp.pot
    with open(path, "w") as f:
        f.write(f"{sp.n} {sp.npt}\n")
        f.write(" ".join(repr(float(x)) for x in sp.box) + "\n")
        f.write(f"{float(sp.rcut)!r} {float(sp.alpha)!r} {float(p['epsi4pi'])!r}\n")
        f.write(f"{sp.nmesh} {sp.order} {int(sp.lsurf)}\n")
        f.write(f"{p['form']} {int(p['lewald'])}\n")
        for k in ("z", "sig", "eps", "rad"):
            f.write(" ".join(repr(float(x)) for x in p[k][: sp.npt]) + "\n")
        for i in range(sp.n):
            x, y, z = (float(v) for v in sp.ro[i])
            f.write(f"{x!r} {y!r} {z!r} {float(sp.az[i])!r} {int(sp.ipt[i]) + 1}\n")
        f.write(f"{len(trials)}\n")
        for ip, r in trials:
            f.write(f"{int(ip) + 1} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")


def run(path):
    return subprocess.run([DRIVER, path], capture_output=True, text=True, timeout=120)


def parse(out):
    res = {"force": {}, "dutotal": [], "dutotal_batch": []}
    for line in out.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "force":
            res["force"][int(t[1]) - 1] = [float(x) for x in t[2:5]]
        elif t[0] in ("dutotal", "dutotal_batch"):
            res[t[0]].append((int(t[2]), float(t[3])))
        elif t[0] == "cells":
            res["cells"] = ([int(x) for x in t[1:4]], int(t[4]))
        elif t[0] in ("u_tot", "u_rec", "u_twob0", "virial", "u_tot_after"):
            res[t[0]] = float(t[1])
        elif t[0] == "mc_pass":
            res["mc_pass"] = (int(t[1]), float(t[2]))
        elif t[0] == "rng_state":
            res["rng_state"] = [int(x) for x in t[1:4]]
    n = len(res["force"])
    res["force"] = np.array([res["force"][i] for i in range(n)])
    return res
