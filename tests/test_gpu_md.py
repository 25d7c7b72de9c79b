"""GPU parity of the Verlet-list path and of velocity-Verlet MD (VelVer, Src/md.F90:252-410)
against the FP64 oracle.

* UTotal over the Verlet list (rc + skin) equals the oracle's cell-list UTotal: forces within
  1e-5 relative (mixed) / 1e-9 (FP64).
* MD: a Python velocity-Verlet loop driven by the oracle's forces (VelVer's update order and
  PBC), is the trajectory reference; after 10 steps the FP64 path's positions agree to 1e-9 A
  and the mixed path's to 1e-5 A.
* Energy conservation over a few hundred steps, and the skin rule triggers rebuilds.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from gpu_helpers import cabi_table, compare, engine, oracle_system, vv_reference

pytestmark = pytest.mark.gpu


def _fluid(n=3000, L=100.0, rcut=10.0, skin=2.0, seed=6):
    sp = systems.small_spme_system(seed=seed, n=n, L=L, nmesh=32, alpha=0.3, rcut=rcut)
    pot = dict(sp.pot, rad=[0.0] * sp.npt)   # LJ repulsion only: no hard core for dynamics
    return systems.SystemSpec(**{**sp.__dict__, "list_mode": "verlet", "drnlist": skin, "pot": pot})


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-9)])
def test_verlet_utotal_matches_oracle(precision, tol):
    sp = _fluid()
    print(f"verlet n={sp.n} precision={precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(systems.SystemSpec(**{**sp.__dict__, "list_mode": "cell"}), tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, tol, tol)
    # displaced positions below the skin: the same list must still give the oracle's result
    rng = np.random.default_rng(1)
    ro2 = sp.ro + rng.uniform(-0.5, 0.5, sp.ro.shape) / np.sqrt(3)
    ro2 = (ro2 + 50.0) % 100.0 - 50.0
    eng.set_positions(ro2)
    got = eng.utotal()
    assert eng.nlist_builds() == 1
    osys2 = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro2, "list_mode": "cell"}), tab)
    compare(got, osys2.utotal(), tol, tol)
    # that pass left the cell order out; trial moves and the cell list complete it for the
    # current positions (trial energies 2e-5 kJ/mol + 1e-6 relative mixed, 1e-8 kJ/mol FP64)
    atol, rtol = (2e-5, 1e-6) if precision == "mixed" else (1e-8, 0.0)
    ips = rng.integers(0, sp.n, size=16)
    rn = (ro2[ips] + rng.uniform(-1.0, 1.0, (16, 3)) + 50.0) % 100.0 - 50.0
    de = eng.dutotal_batch(ips, rn)
    for t in range(len(ips)):
        ref = osys2.dutotal(int(ips[t]), rn[t])
        assert de["overlap"][t] == ref["overlap"]
        if not ref["overlap"]:
            assert abs(de["du_tot"][t] - ref["du_tot"]) <= atol + rtol * abs(ref["du_tot"]), t
    nc, start, cip = eng.cell_list()
    onc, ocs, _ = oracle.cell_init(sp.box, sp.rcut + sp.drnlist, sp.rcut)
    head, nxt, npart, _ = oracle.cell_build(ro2, sp.box, onc, ocs)
    assert np.array_equal(nc, onc) and np.array_equal(np.diff(start), npart)
    for c, m in enumerate(oracle.cell_members(head, nxt)):
        assert list(cip[start[c]:start[c + 1]]) == m
    eng.close()


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-9)])
def test_verlet_dense_list_matches_oracle(precision, tol):
    """A dense list (about 550 partners per particle, list width 1152): too wide for the build
    to stage in LDS, so its partners go straight to the list; one collapsed cell."""
    sp = systems.small_spme_system(seed=8, n=1200, L=44.0, nmesh=32, alpha=0.22, rcut=18.0)
    sp = systems.SystemSpec(**{**sp.__dict__, "list_mode": "verlet", "drnlist": 3.0})
    print(f"dense verlet n={sp.n} precision={precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    osp = systems.SystemSpec(**{**sp.__dict__, "list_mode": "cell"})
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), oracle_system(osp, tab).utotal(), tol, tol)
    rng = np.random.default_rng(2)
    ro2 = sp.ro + rng.uniform(-1.0, 1.0, sp.ro.shape) / np.sqrt(3)
    ro2 = (ro2 + 22.0) % 44.0 - 22.0
    eng.set_positions(ro2)
    got = eng.utotal()
    assert eng.nlist_builds() == 1
    compare(got, oracle_system(systems.SystemSpec(**{**osp.__dict__, "ro": ro2}), tab).utotal(), tol, tol)
    eng.close()


@pytest.mark.parametrize("precision,xtol", [("fp64", 1e-9), ("mixed", 1e-5)])
def test_md_trajectory_matches_oracle(precision, xtol):
    sp = _fluid(n=1500, L=80.0)
    tab = oracle.build_table(sp.pot)
    mass = [20.0, 35.0]
    rng = np.random.default_rng(3)
    v0 = rng.normal(size=sp.ro.shape) * 3.0
    dt, nsteps, scale = 0.002, 10, 100.0
    print(f"md n={sp.n} precision={precision}", flush=True)
    from molsim_amd.engine import Engine
    eng = Engine(sp, precision=precision, table=cabi_table(tab), force_scale=scale)
    eng.md_init(mass, v0)
    log = eng.md_run(nsteps, dt)
    x = eng.positions()
    v = eng.velocities()
    xr, vr, _, _ = vv_reference(sp, tab, mass, v0, dt, nsteps, scale)
    dx = x - xr
    dx -= sp.box[0] * np.rint(dx / sp.box[0])
    print(f"max |dx| {np.abs(dx).max():.2e}, max |dv| {np.abs(v - vr).max():.2e}", flush=True)
    assert np.abs(dx).max() < xtol
    assert np.abs(v - vr).max() < xtol * 1e3
    assert np.all(np.isfinite(log))
    eng.close()


def test_md_energy_conservation_and_rebuilds():
    sp = _fluid(n=4000, L=110.0, skin=1.0)
    from molsim_amd.engine import Engine
    eng = Engine(sp, precision="mixed", force_scale=100.0)
    rng = np.random.default_rng(8)
    mass = [20.0, 35.0]
    minv = (1.0 / np.asarray(mass))[sp.ipt][:, None]
    v0 = rng.normal(size=sp.ro.shape) * np.sqrt(2.479 * 100.0 * minv)  # ~kT per dof at 298 K
    eng.md_init(mass, v0)
    dt = 0.004
    log = eng.md_run(300, dt)
    etot = log[:, 0] * 100.0 + log[:, 1]          # U in force-scaled units (kJ/mol * 100) + K
    drift = np.abs(etot - etot[0]).max() / abs(log[0, 1])
    print(f"energy drift {drift:.2e} of K over 300 steps; list builds {eng.nlist_builds()}", flush=True)
    assert drift < 2e-2
    assert eng.nlist_builds() >= 2
    eng.close()
