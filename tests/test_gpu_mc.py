"""GPU parity of the single-particle MC caller (MCPass / SPartMove / Metropolis / MCUpdate)
against the same loop driven by the oracle's DUTotal: with the FP64 path the accept/reject
sequence and the final positions are identical; with the mixed path at least 99 % of the
decisions agree (its SPME potential mesh is FP32)."""
import math

import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import Engine, Random
from gpu_helpers import cabi_table, oracle_system

pytestmark = pytest.mark.gpu


def _oracle_mc(sp, tab, ntrial, dtran, beta, seed):
    rng = oracle.Rng(seed)
    ro = sp.ro.copy()
    L = sp.box[0]
    events = []
    # MOLSIM's linked cell lists, updated by UpdateCellIp on every accepted move
    cells = oracle.LinkedCellList(sp.ro, sp.box, sp.rcut + sp.drnlist, sp.rcut) if sp.list_mode == "cell" else None
    osys = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro}), tab)
    nxt = 0
    for _ in range(ntrial):
        ip = nxt
        nxt = (nxt + 1) % sp.n
        rn = ro[ip] + np.array([(rng() - 0.5) * dtran for _ in range(3)])
        for d in range(3):
            if abs(rn[d]) > L / 2:
                rn[d] -= L * np.round(rn[d] / L)
        osys.set_positions(ro)
        du = osys.dutotal(ip, rn)
        if du["overlap"]:
            acc = False
        else:
            dured = beta * du["du_tot"]
            if dured > 87.0:
                acc = False
            elif dured < -87.0:
                acc = True
            else:
                fac = math.exp(-dured)
                acc = True if fac > 1.0 else fac > rng()
        events.append(-1 if du["overlap"] else int(acc))
        if acc:
            ro[ip] = rn
            if cells is not None:
                cells.update(ip, rn)
    return np.array(events), ro, rng, cells


@pytest.mark.parametrize("ewald,precision", [("std", "fp64"), ("std_cell", "fp64"), ("spm", "fp64"),
                                             ("spm", "mixed"), ("spm_k40_p5", "fp64"),
                                             ("spm_all_p11", "fp64")])
def test_mc_pass_matches_oracle_loop(ewald, precision):
    """std: the device-resident chain (all pairs); std_cell: the device chain on a 3x3x3 cell
    grid (accepted moves into another cell patch the cell list inside the launch, as
    UpdateCellIp does); spm: the device chain with SPME, phi corrected inside the launch for
    the moves accepted since the last UTotal pass and recomputed after every 16 (15 launches
    for the 228 accepted moves of this pass); spm_k40_p5: the reference decks' mesh (40^3, order 5, the general
    mesh path); spm_all_p11: all pairs, order 11.  With a cell list, the cell order after the
    pass equals MOLSIM's linked lists after the same UpdateCellIp calls."""
    sp = systems.small_spme_system(n=200, L=36.0, nmesh=32, alpha=0.3, rcut=11.0)
    if ewald == "std":
        sp = systems.SystemSpec(**{**sp.__dict__, "ewald": "std", "ncut": 6, "list_mode": "all",
                                   "rcut": 17.9})
        sp.pot["rcut"] = 17.9
    elif ewald == "std_cell":
        sp = systems.SystemSpec(**{**sp.__dict__, "ewald": "std", "ncut": 6, "list_mode": "cell"})
    elif ewald == "spm_k40_p5":
        sp = systems.small_spme_system(n=200, L=36.0, nmesh=40, order=5, alpha=0.3, rcut=11.0)
    elif ewald == "spm_all_p11":
        sp = systems.small_spme_system(n=200, L=36.0, nmesh=32, order=11, alpha=0.3, rcut=17.9,
                                       list_mode="all")
    tab = oracle.build_table(sp.pot)
    ntrial, dtran, beta = 300, 2.0, systems.beta_kjmol()
    print(f"mc ewald={ewald} precision={precision} n={sp.n}", flush=True)
    eng = Engine(sp, precision=precision, table=cabi_table(tab))
    rng = Random(-4711)
    dsum, nacc, ev = eng.mc_pass(ntrial, dtran, beta, rng)
    ref_ev, ref_ro, ref_rng, ref_cells = _oracle_mc(sp, tab, ntrial, dtran, beta, -4711)
    agree = float(np.mean(ev == ref_ev))
    print(f"accepted {nacc}/{ntrial}, decisions agreeing with the oracle loop {agree:.4f}", flush=True)
    if precision == "fp64":
        assert np.array_equal(ev, ref_ev)
        ro = eng.positions()
        assert np.abs(ro - ref_ro).max() < 1e-9
        # the generator state carried through the pass continues the reference's sequence
        assert rng() == ref_rng()
        if ref_cells is not None:
            nc, start, cip = eng.cell_list()
            got = [list(cip[start[c]:start[c + 1]]) for c in range(len(start) - 1)]
            assert got == ref_cells.members()
        # and the state after the pass equals a fresh evaluation of the final positions
        full = eng.utotal()
        ref = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ref_ro}), tab).utotal()
        assert abs(full["u_tot"] - ref["u_tot"]) <= 1e-9 * abs(ref["u_tot"]) + 1e-9
    else:
        assert agree >= 0.99
    assert 0 < nacc < ntrial
    eng.close()


@pytest.mark.parametrize("ewald", ["std", "spm"])
def test_mc_chain_resumes_from_cnf(ewald, tmp_path):
    """A chain interrupted after 150 trials, saved as a MOLSIM .cnf (positions and the
    generator state iseed/am/ix/iy, Src/molsim.F90:936-938) and resumed in a fresh context,
    makes the same decisions and ends at the same positions as the uninterrupted chain."""
    from molsim_amd import cnf
    sp = systems.small_spme_system(n=200, L=36.0, nmesh=32, alpha=0.3, rcut=11.0)
    if ewald == "std":
        sp = systems.SystemSpec(**{**sp.__dict__, "ewald": "std", "ncut": 6, "list_mode": "cell"})
    tab = cabi_table(oracle.build_table(sp.pot))
    dtran, beta = 2.0, systems.beta_kjmol()
    print(f"mc resume ewald={ewald} n={sp.n}", flush=True)
    a = Engine(sp, precision="fp64", table=tab)
    ev_all = a.mc_pass(300, dtran, beta, Random(-4711), isamp=1)[2]
    ro_all = a.positions()
    a.close()
    b = Engine(sp, precision="fp64", table=tab)
    g = Random(-4711)
    ev1 = b.mc_pass(150, dtran, beta, g, isamp=1)[2]
    path = tmp_path / "mid.cnf"
    cnf.write_cnf(path, cnf.from_engine(b, istep=1, rng=g))
    b.close()
    c = cnf.read_cnf(path)
    sp2 = cnf.apply_to_spec(sp, c)
    e = Engine(sp2, precision="fp64", table=tab)
    ev2 = e.mc_pass(150, dtran, beta, Random.from_state(*c.rng_state()), isamp=1)[2]
    assert np.array_equal(np.concatenate([ev1, ev2]), ev_all)
    assert np.abs(e.positions() - ro_all).max() < 1e-9
    e.close()
