"""GPU parity of DUTotal (single-particle MC trials, all FP64) and of the standard-Ewald
UTotal against the FP64 oracle, which follows the reference's own algorithms (the SPME
trial energy from a full r2c of the modified mesh, the std-Ewald one from updated
structure factors).

Tolerance of a trial energy: |dE - dE_oracle| <= 2e-5 kJ/mol + 1e-6 |dE_oracle| on the
mixed path (its SPME potential mesh comes from the FP32 FFT), 1e-8 kJ/mol on the FP64 path.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import Engine
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu


def _trials(sp, n, seed, dmax=1.0):
    rng = np.random.default_rng(seed)
    ips = rng.integers(0, sp.n, size=n)
    L = sp.box[0]
    rn = sp.ro[ips] + rng.uniform(-dmax, dmax, size=(n, 3))
    rn = (rn + L / 2) % L - L / 2
    return ips, rn


@pytest.mark.parametrize("precision,atol,rtol", [("mixed", 2e-5, 1e-6), ("fp64", 1e-8, 0.0)])
@pytest.mark.parametrize("ewald", ["spm", "std", "std_cube"])
def test_dutotal_trials(ewald, precision, atol, rtol):
    """std_cube: the k-vectors of ncutregion = 'cube' (n^2 <= 3 ncut^2, Src/energy.F90:6754-6760)."""
    sp = systems.small_spme_system(n=300, L=40.0, nmesh=32, alpha=0.3, rcut=12.0)
    if ewald == "std":
        sp = systems.SystemSpec(**{**sp.__dict__, "ewald": "std", "ncut": 6})
    elif ewald == "std_cube":
        sp = systems.SystemSpec(**{**sp.__dict__, "ewald": "std", "ncut": 5, "ncutregion": "cube"})
    print(f"dutotal ewald={ewald} precision={precision} n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    osys = oracle_system(sp, tab)
    eng = engine(sp, tab, precision)
    ips, rn = _trials(sp, 24, 11)
    got = eng.dutotal_batch(ips, rn)
    nchk = 0
    for t in range(len(ips)):
        ref = osys.dutotal(int(ips[t]), rn[t])
        assert got["overlap"][t] == ref["overlap"]
        if ref["overlap"]:
            continue
        err = abs(got["du_tot"][t] - ref["du_tot"])
        assert err <= atol + rtol * abs(ref["du_tot"]), (t, got["du_tot"][t], ref["du_tot"])
        assert abs(got["du_rec"][t] - ref["du_rec"]) <= atol + rtol * abs(ref["du_rec"])
        nchk += 1
    assert nchk >= 12
    # single-trial entry point with per-type-pair output
    one = eng.dutotal(int(ips[0]), rn[0])
    ref = osys.dutotal(int(ips[0]), rn[0])
    if not ref["overlap"]:
        assert np.allclose(one["du_twob"], ref["du_twob"], atol=atol * 10)
    eng.close()


def test_mc_commit_sequence():
    """Accept a few moves (MCUpdate) and check later trials against the oracle's state."""
    sp = systems.small_spme_system(n=250, L=40.0, nmesh=32, alpha=0.3, rcut=12.0)
    tab = oracle.build_table(sp.pot)
    eng = engine(sp, tab)
    ro = sp.ro.copy()
    rng = np.random.default_rng(5)
    accepted = 0
    for step in range(30):
        ip = int(rng.integers(sp.n))
        rn = (ro[ip] + rng.uniform(-1, 1, 3) + 20.0) % 40.0 - 20.0
        d = eng.dutotal(ip, rn)
        cur = systems.SystemSpec(**{**sp.__dict__, "ro": ro})
        ref = oracle_system(cur, tab).dutotal(ip, rn)
        assert d["overlap"] == ref["overlap"]
        if d["overlap"]:
            continue
        assert abs(d["du_tot"] - ref["du_tot"]) <= 2e-5 + 1e-6 * abs(ref["du_tot"])
        if d["du_tot"] < 0.5:
            eng.mc_update(ip, rn)
            ro[ip] = rn
            accepted += 1
    assert accepted >= 5
    # the accumulated state equals a fresh evaluation
    full = eng.utotal()
    ref = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro}), tab).utotal()
    compare(full, ref, 1e-5, 1e-5)
    eng.close()


@pytest.mark.parametrize("region", ["sphere", "cube"])
@pytest.mark.parametrize("precision,ftol,etol", [("mixed", 1e-5, 1e-5), ("fp64", 1e-9, 1e-9)])
def test_std_ewald_electrolyte(precision, ftol, etol, region):
    """The 1000-ion primitive-model electrolyte: std Ewald (|n| <= 10), rc = L/2, all pairs;
    k-vectors in the sphere n^2 <= 100 or the cube n^2 <= 300 (1330 octant vectors)."""
    sp = systems.electrolyte_primitive()
    sp = systems.SystemSpec(**{**sp.__dict__, "ncutregion": region})
    print(f"electrolyte n={sp.n} precision={precision} region={region}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, ftol, etol)
    eng.close()
