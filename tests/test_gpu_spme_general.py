"""SPME at every mesh size and order the reference accepts (EwaldSetup takes any even nmesh,
Src/energy.F90:6891-6892; nmlPotential's default nmesh = 48 and order 5, Src/potential.F90:359-360;
the test decks' nmesh = 40, Testin/in/*.ewspm.in; order 11 in TestEwaldSPM, :3268), against the
FP64 oracle: forces 1e-5 relative (mixed) / 1e-9 (FP64), energies likewise against the energy
scale.

* the reference's defaults (K = 48, order 5) and its decks' K = 40 with order 6;
* meshes whose sizes have factors 3, 5, 7, 11 and orders 2 and 8..11;
* the geometry of Testin/in/mic.mc.ewspm.in (20 macroions of charge -60, radius 20 A, 1200
  counterions, box 342 A, rcut 170 A, nmesh 40, order 6, no surface term);
* MC trial energies at K = 48, order 5, and a batched pass holding a general-mesh system.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu


def _sys(K, order, n=1500, L=70.0, alpha=0.32, rcut=11.0, seed=21):
    sp = systems.small_spme_system(seed=seed, n=n, L=L, nmesh=K, order=order, alpha=alpha, rcut=rcut)
    return sp


@pytest.mark.parametrize("K,order", [(48, 5), (40, 6), (40, 5), (48, 6), (54, 6), (50, 4), (42, 7),
                                     (22, 3), (36, 8), (40, 9), (40, 11), (30, 2)])
@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-9)])
def test_general_mesh_matches_oracle(K, order, precision, tol):
    sp = _sys(K, order)
    print(f"general mesh K={K} order={order} {precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, tol, tol)
    compare(eng.utotal(), ref, tol, tol)  # again (a graph replay on the mixed path)
    eng.close()


def _alpha_from_tolerance(utol, q2sum, epsi4pi, L, rcut):
    """RcutToAlpha (Src/energy.F90:7090-7105, lq2sum = 0): the deck's alpha for its tolerance."""
    y = (utol / (q2sum * epsi4pi)) * np.sqrt(2.0 * (L * rcut) ** 3)
    x = 0.1
    for _ in range(40):
        xn = (1.0 / rcut) * np.sqrt(-np.log(y * x ** 2))
        if abs(x - xn) / xn < 1e-6:
            return xn
        x = xn
    raise RuntimeError("no convergence")


def _micelle_deck(seed=5):
    """Testin/in/mic.mc.ewspm.in at its last boxlen (342 A): 20 macroions (z = -60, radat = 20)
    and 1200 counterions (z = 1, radat = 2), placed at random without overlap."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    L = 342.0
    mac = []
    while len(mac) < 20:
        p = rng.uniform(0.0, L, 3)
        if all(np.linalg.norm((p - q + L / 2) % L - L / 2) >= 40.0 for q in mac):
            mac.append(p)
    mac = np.array(mac)
    tm = cKDTree(mac, boxsize=L)
    ions = rng.uniform(0.0, L, (1200, 3))
    for _ in range(200):
        bad = tm.query_ball_point(ions, 22.0, return_length=True) > 0
        pairs = cKDTree(ions, boxsize=L).query_pairs(4.0, output_type="ndarray")
        if len(pairs):
            bad[pairs[:, 1]] = True
        if not bad.any():
            break
        ions[bad] = rng.uniform(0.0, L, (int(bad.sum()), 3))
    ro = np.concatenate([mac, ions]) - L / 2
    ipt = np.array([0] * 20 + [1] * 1200, dtype=np.int32)
    az = np.where(ipt == 0, -60.0, 1.0)
    relp = 78.4026
    epsi4pi = systems.EPSI0FOURPI / relp
    rcut = 170.0
    alpha = _alpha_from_tolerance(1e-5, float((az * az).sum()), epsi4pi, L, rcut)
    pot = systems._pot(2, [-60.0, 1.0], 0, rcut, relp, alpha, rad=[20.0, 2.0])
    return systems.SystemSpec("micelle_deck", ro, ipt, az, np.array([L, L, L]), pot, rcut,
                              drnlist=30.0, list_mode="cell", ewald="spm", alpha=alpha, nmesh=40,
                              order=6, lsurf=False)


@pytest.mark.parametrize("precision,tol", [("mixed", 1e-5), ("fp64", 1e-9)])
def test_micelle_deck_geometry(precision, tol):
    sp = _micelle_deck()
    print(f"micelle deck alpha={sp.alpha:.5f} {precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, tol, tol)
    eng.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_general_mesh_trials(precision):
    sp = _sys(48, 5, seed=22)
    tab = oracle.build_table(sp.pot)
    osys = oracle_system(sp, tab)
    eng = engine(sp, tab, precision)
    eng.utotal()
    rng = np.random.default_rng(3)
    ips = rng.integers(0, sp.n, size=24)
    rn = (sp.ro[ips] + rng.uniform(-1.5, 1.5, (24, 3)) + 35.0) % 70.0 - 35.0
    de = eng.dutotal_batch(ips, rn)
    atol, rtol = (2e-5, 1e-6) if precision == "mixed" else (1e-8, 0.0)
    nchk = 0
    for t in range(len(ips)):
        ref = osys.dutotal(int(ips[t]), rn[t])
        assert de["overlap"][t] == ref["overlap"]
        if not ref["overlap"]:
            nchk += 1
            assert abs(de["du_tot"][t] - ref["du_tot"]) <= atol + rtol * abs(ref["du_tot"]), t
    assert nchk > 10
    eng.close()


def test_general_mesh_in_a_batch():
    """A batched pass over a power-of-two mesh system and two general ones equals each system
    alone, bit for bit."""
    from molsim_amd.engine import Batch
    specs = [_sys(32, 6, n=800, L=60.0, seed=31), _sys(40, 6, n=900, L=62.0, seed=32),
             _sys(48, 5, n=1000, L=64.0, seed=33)]
    tabs = [oracle.build_table(sp.pot) for sp in specs]
    alone = []
    for sp, tab in zip(specs, tabs):
        e = engine(sp, tab, "mixed")
        alone.append(e.utotal())
        e.close()
    engs = [engine(sp, tab, "mixed") for sp, tab in zip(specs, tabs)]
    b = Batch(engs)
    for rep in range(2):
        got = b.utotal()
        for i in range(len(engs)):
            assert got[i]["u_tot"] == alone[i]["u_tot"], (rep, i)
            assert got[i]["u_rec"] == alone[i]["u_rec"], (rep, i)
            assert np.array_equal(got[i]["force"], alone[i]["force"]), (rep, i)
    b.close()
    for e in engs:
        e.close()
