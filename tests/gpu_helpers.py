"""Shared helpers of the GPU parity tests."""
import numpy as np

import oracle
from molsim_amd import systems
from molsim_amd.engine import Engine

LIST = {"all": 0, "cell": 1}
EWALD = {"none": 0, "std": 1, "spm": 2}


def cabi_table(tab):
    """Oracle table (0-based record offsets) in the C-ABI's MOLSIM convention (1-based)."""
    t = dict(tab)
    t["iubuflow"] = (np.asarray(tab["iubuflow"]) + 1).astype(np.int32)
    return t


def engine(sp, tab, precision="mixed"):
    return Engine(sp, precision=precision, table=cabi_table(tab))


def oracle_system(sp, tab):
    """The oracle's system of a spec.  A Verlet-list spec is evaluated over all pairs (the same
    pairs, r^2 <= rcut^2); weak charges (sp.laz) carry over with az = z[type] or 0."""
    laz = getattr(sp, "laz", None)
    zat = None if laz is None else np.asarray(sp.pot["z"][: sp.npt], dtype=np.float64)
    az = sp.az if laz is None else np.where(np.asarray(laz) != 0, zat[sp.ipt], 0.0)
    return oracle.System(sp.ro, sp.ipt, az, sp.box, sp.rcut, sp.pot["epsi4pi"], tab, npt=sp.npt,
                         list_mode=LIST.get(sp.list_mode, 0), drnlist=sp.drnlist, ewald=EWALD[sp.ewald],
                         surf=sp.lsurf, alpha=sp.alpha, ncut=sp.ncut,
                         region={"sphere": 0, "cube": 1}[sp.ncutregion], K=sp.nmesh, order=sp.order,
                         laz=laz, zat=zat)


def rel_force_err(f, ref):
    return float(np.linalg.norm(f - ref) / np.linalg.norm(ref))


def compare(gpu, ref, ftol, etol):
    """Relative errors of forces (2-norm over all particles) and energies against the oracle."""
    ferr = rel_force_err(gpu["force"], ref["force"])
    scale = abs(ref["u_tot"]) + abs(ref["u_rec"]) + np.abs(ref["u_twob"]).max()
    errs = dict(force=ferr,
                u_tot=abs(gpu["u_tot"] - ref["u_tot"]) / scale,
                u_rec=abs(gpu["u_rec"] - ref["u_rec"]) / scale,
                u_twob=float(np.abs(gpu["u_twob"] - ref["u_twob"]).max() / scale),
                virial=abs(gpu["virial"] - ref["virial"]) / (abs(ref["virial"]) + scale))
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert ferr <= ftol, errs
    for k in ("u_tot", "u_rec", "u_twob", "virial"):
        assert errs[k] <= etol, errs
    return errs


def anint(x):
    """Fortran ANINT: nearest integer, halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def vv_reference(sp, tab, mass, v0, dt, nsteps, scale, inlist=0):
    """VelVer (Src/md.F90:252-410) + NList(iSimulationStep) (Src/nlist.F90:597-621) driven by
    oracle forces: drostep = dt v + (dt^2/2) a, ro += drostep, PBC with ANINT, then
    a = F / m and v += (dt/2)(a + a_old); after every step the displacement sums decide a
    rebuild by the oracle's skin test (inlist = 0) or every inlist-th step.  Returns positions,
    velocities, the rebuild steps (1-based) and the positions of the last rebuild."""
    L = np.asarray(sp.box, dtype=np.float64)
    osp = systems.SystemSpec(**{**sp.__dict__, "list_mode": "cell"})
    x = sp.ro.copy()
    v = v0.copy()
    minv = (1.0 / np.asarray(mass))[sp.ipt][:, None]
    a = minv * (oracle_system(osp, tab).utotal()["force"] * scale)
    drosum = np.zeros_like(x)
    fac = 0.5 * (dt * dt)
    rebuilds, x_build = [], x.copy()
    for step in range(1, nsteps + 1):
        ds = dt * v + fac * a
        x = x + ds
        x = np.where(np.abs(x) > L / 2, x - L * anint(x * (1.0 / L)), x)
        drosum += ds
        f = oracle_system(systems.SystemSpec(**{**osp.__dict__, "ro": x}), tab).utotal()["force"]
        an = minv * (f * scale)
        v = v + (0.5 * dt) * (an + a)
        a = an
        rb = (step % inlist == 0) if inlist > 0 else oracle.nlist_skin(drosum, sp.drnlist)[0]
        if rb:
            drosum[:] = 0.0
            rebuilds.append(step)
            x_build = x.copy()
    return x, v, rebuilds, x_build
