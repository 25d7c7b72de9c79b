"""UpdateCellIp on the device (Src/celllist.F90:329-382, called by MCUpdate, Src/mc.F90:5232):
after a sequence of accepted moves (msg_commit) the cell list equals MOLSIM's linked lists
after the same UpdateCellIp calls, bit for bit (membership and order: a moved particle becomes
its new cell's head); the energies and trial energies of the patched state equal the oracle's
at the final positions.  Standard Ewald keeps its structure factors and needs no pass between
moves; SPME refreshes its mesh by a pass that keeps the patched order (including the
wave-per-cell-chunk pair tile)."""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu


def _moves(sp, n, seed, step=6.0):
    rng = np.random.default_rng(seed)
    ips = rng.integers(0, sp.n, size=n)
    L = sp.box
    return ips, rng.uniform(-step, step, (n, 3)), L


@pytest.mark.parametrize("ewald,precision,tile", [("std", "fp64", None), ("std", "mixed", None),
                                                  ("spm", "mixed", None), ("spm", "fp64", None),
                                                  ("spm", "mixed", 0), ("none", "mixed", None)])
def test_commits_patch_cell_list(ewald, precision, tile):
    sp = systems.small_spme_system(seed=41, n=1500, L=72.0, nmesh=32, alpha=0.3, rcut=11.0)
    pot = dict(sp.pot, rad=[0.0] * sp.npt)   # arbitrary moves: no hard core
    kw = dict(pot=pot, ewald=ewald)
    if ewald == "std":
        kw["ncut"] = 6
    sp = systems.SystemSpec(**{**sp.__dict__, **kw})
    tab = oracle.build_table(sp.pot)
    print(f"commits ewald={ewald} {precision} tile={tile}", flush=True)
    eng = engine(sp, tab, precision)
    if tile is not None:
        eng.set_tile(tile)
    eng.utotal()
    cells = oracle.LinkedCellList(sp.ro, sp.box, sp.rcut + sp.drnlist, sp.rcut)
    ro = sp.ro.copy()
    ips, steps, L = _moves(sp, 120, 7)
    ncross = 0
    for k, (ip, dr) in enumerate(zip(ips, steps)):
        rn = (ro[ip] + dr + 0.5 * L) % L - 0.5 * L
        # a trial first (the state must be current); moves onto another particle are not
        # committed, as MC rejects them
        if eng.dutotal(int(ip), rn)["overlap"]:
            continue
        eng.mc_update(int(ip), rn)
        ro[ip] = rn
        ncross += cells.update(int(ip), rn)
        if k % 40 == 39:
            nc, start, cip = eng.cell_list()
            got = [list(cip[start[c]:start[c + 1]]) for c in range(len(start) - 1)]
            assert got == cells.members(), k
    print(f"{ncross} of {len(ips)} moves changed cell", flush=True)
    assert ncross > 30
    osys = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro}), tab)
    tol = 1e-5 if precision == "mixed" else 1e-9
    compare(eng.utotal(), osys.utotal(), tol, tol)
    # the pass kept the patched order
    nc, start, cip = eng.cell_list()
    assert [list(cip[start[c]:start[c + 1]]) for c in range(len(start) - 1)] == cells.members()
    # trial energies of the patched state
    rng = np.random.default_rng(8)
    tips = rng.integers(0, sp.n, size=8)
    trn = (ro[tips] + rng.uniform(-1.0, 1.0, (8, 3)) + 0.5 * L) % L - 0.5 * L
    de = eng.dutotal_batch(tips, trn)
    atol, rtol = (2e-5, 1e-6) if precision == "mixed" else (1e-8, 0.0)
    for t in range(8):
        ref = osys.dutotal(int(tips[t]), trn[t])
        assert abs(de["du_tot"][t] - ref["du_tot"]) <= atol + rtol * abs(ref["du_tot"]), t
    # new positions from the caller: the next pass sorts the cells afresh (SetCellList)
    eng.set_positions(ro)
    eng.utotal()
    fresh = oracle.LinkedCellList(ro, sp.box, sp.rcut + sp.drnlist, sp.rcut)
    nc, start, cip = eng.cell_list()
    assert [list(cip[start[c]:start[c + 1]]) for c in range(len(start) - 1)] == fresh.members()
    eng.close()
