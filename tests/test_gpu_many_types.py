"""Systems of 5..8 particle types (MOLSIM sizes u%twob(0:nptpt) at run time, Src/mol.F90:92-115;
three of its test decks use 5 to 7 types): the pair kernel that reads the table of up to 36
type pairs from memory, against the FP64 oracle.

* UTotal with a cell list and SPME (mixed and FP64 paths), with a Verlet list, and over all
  pairs with standard Ewald: per-type-pair energies u_twob, forces 1e-5 relative (mixed) /
  1e-9 (FP64);
* DUTotal trial energies and a short velocity-Verlet run (FP64) against an oracle-driven loop;
* eight types (the C-ABI's limit) work, nine are refused, and a batch refuses such systems.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import Batch, Engine, MolsimGpuError
from gpu_helpers import cabi_table, compare, engine, oracle_system, vv_reference

pytestmark = pytest.mark.gpu

CHARGES6 = (1.0, -1.0, 2.0, -2.0, 1.0, -1.0)


def _many(nt=6, n=1200, L=56.0, rcut=12.0, list_mode="cell", seed=31, **kw):
    charges = (CHARGES6 + (1.0, -1.0))[:nt]
    sp = systems.small_spme_system(seed=seed, n=n, L=L, nmesh=32, alpha=0.3, rcut=rcut,
                                   charges=charges, list_mode=list_mode)
    pot = dict(sp.pot)
    # a different LJ diameter and depth per type, so every type pair has a table of its own
    pot["sig"] = [3.0 + 0.1 * t for t in range(nt)]
    pot["eps"] = [e * (1.0 + 0.2 * t) for t, e in enumerate(pot["eps"])]
    pot["rad"] = [0.0] * nt
    return systems.SystemSpec(**{**sp.__dict__, "pot": pot, **kw})


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-9)])
def test_cell_list_spme(precision, tol):
    sp = _many()
    print(f"6 types cell/SPME n={sp.n} {precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    got = eng.utotal()
    assert got["u_twob"].size == 6 * 7 // 2 + 1
    compare(got, ref, tol, tol)
    compare(eng.utotal(), ref, tol, tol)
    eng.close()


def test_verlet_list_and_md():
    sp = _many(list_mode="verlet", drnlist=1.0, seed=32)
    print(f"6 types Verlet n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    osp = systems.SystemSpec(**{**sp.__dict__, "list_mode": "cell"})
    compare(engine(sp, tab, "mixed").utotal(), oracle_system(osp, tab).utotal(), 1e-5, 1e-5)
    mass = [20.0 + 5.0 * t for t in range(sp.npt)]
    rng = np.random.default_rng(3)
    v0 = rng.normal(size=sp.ro.shape) * 2.0
    eng = Engine(sp, precision="fp64", table=cabi_table(tab), force_scale=100.0)
    eng.md_init(mass, v0)
    eng.md_run(8, 0.002)
    xr, _, _, _ = vv_reference(sp, tab, mass, v0, 0.002, 8, 100.0)
    dx = eng.positions() - xr
    dx -= sp.box[0] * np.rint(dx / sp.box[0])
    print(f"max |dx| {np.abs(dx).max():.2e}", flush=True)
    assert np.abs(dx).max() < 1e-9
    eng.close()


def test_all_pairs_std_ewald_and_trials():
    sp = _many(n=600, L=40.0, rcut=20.0, list_mode="all", ewald="std", alpha=0.15, ncut=6, nmesh=0,
               order=0, seed=33)
    print(f"6 types all pairs / std Ewald n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    osys = oracle_system(sp, tab)
    eng = engine(sp, tab, "fp64")
    compare(eng.utotal(), osys.utotal(), 1e-9, 1e-9)
    rng = np.random.default_rng(4)
    ips = rng.integers(0, sp.n, size=16)
    rn = (sp.ro[ips] + rng.uniform(-1.0, 1.0, (16, 3)) + 20.0) % 40.0 - 20.0
    de = eng.dutotal_batch(ips, rn)
    nchk = 0
    for t in range(16):
        ref = osys.dutotal(int(ips[t]), rn[t])
        assert de["overlap"][t] == ref["overlap"]
        if not ref["overlap"]:
            nchk += 1
            assert abs(de["du_tot"][t] - ref["du_tot"]) <= 1e-8
            assert abs(de["du_rec"][t] - ref["du_rec"]) <= 1e-8
    assert nchk > 8
    eng.close()


def test_type_limit_and_batch():
    sp8 = _many(nt=8, n=800, L=50.0, seed=34)
    tab = oracle.build_table(sp8.pot)
    compare(engine(sp8, tab, "mixed").utotal(), oracle_system(sp8, tab).utotal(), 1e-5, 1e-5)
    sp9 = systems.small_spme_system(seed=35, n=900, L=50.0, nmesh=32, charges=(1.0, -1.0) * 4 + (0.0,))
    with pytest.raises(MolsimGpuError):
        Engine(sp9, precision="mixed", table=cabi_table(tab))  # refused at creation
    e1, e2 = engine(sp8, tab, "mixed"), engine(_many(seed=36), oracle.build_table(_many(seed=36).pot), "mixed")
    with pytest.raises(MolsimGpuError):
        Batch([e1, e2])
    e1.close()
    e2.close()
