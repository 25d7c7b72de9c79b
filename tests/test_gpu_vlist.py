"""GPU Verlet lists against the oracle, bit for bit (SetVListMD / SetVListMC,
Src/nlist.F90:724-884), and the NList skin test and rebuild timing (Src/nlist.F90:597-621).

* The exported list (msg_get_verlet_list) equals the oracle's rows exactly, for the MD parity
  half list and the MC full list, including a grid with a collapsed dimension and after a
  rebuild at new positions.
* The skin test uses the reference's r2max1 / r2max2 update order: two displacement patterns
  with the same two largest values, one rebuilding and one not, as the oracle says.
* MD (FP64 path): the rebuild steps and the list after the run equal those of an oracle-driven
  velocity-Verlet loop with NList's skin test after every step; inlist > 0 rebuilds every
  inlist-th step of the macrostep; an overflow during a run is reported at the end of the run.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import Engine, MolsimGpuError
from gpu_helpers import cabi_table, engine, oracle_system, vv_reference

pytestmark = pytest.mark.gpu


def _verlet(n=3000, L=100.0, rcut=10.0, skin=2.0, seed=6, box=None):
    sp = systems.small_spme_system(seed=seed, n=n, L=L, nmesh=32, alpha=0.3, rcut=rcut)
    pot = dict(sp.pot, rad=[0.0] * sp.npt)   # no hard core: dynamics and arbitrary moves
    spec = systems.SystemSpec(**{**sp.__dict__, "list_mode": "verlet", "drnlist": skin, "pot": pot})
    if box is not None:  # the same ions inserted uniformly (3 A apart) into a box of other shape
        from scipy.spatial import cKDTree
        box = np.asarray(box, dtype=np.float64)
        rng = np.random.default_rng(seed)
        pts = rng.uniform(0.0, 1.0, (n, 3)) * box
        for _ in range(200):
            pairs = cKDTree(pts, boxsize=box).query_pairs(3.0, output_type="ndarray")
            if not len(pairs):
                break
            bad = np.unique(pairs[:, 1])
            pts[bad] = rng.uniform(0.0, 1.0, (len(bad), 3)) * box
        spec = systems.SystemSpec(**{**spec.__dict__, "ro": pts - 0.5 * box, "box": box})
    return spec


def _check_list(eng, sp, kind, maxn=400):
    nn, jl = eng.verlet_list(kind, maxn)
    onn, ojl = oracle.vlist(kind, sp.ro, sp.box, sp.rcut, sp.drnlist, maxn=maxn, table=True)
    print(f"{kind}: {int(nn.sum())} entries, longest row {int(nn.max())}", flush=True)
    assert np.array_equal(nn, onn)
    assert np.array_equal(jl, ojl)


@pytest.mark.parametrize("kind", ["md", "mc"])
@pytest.mark.parametrize("box", [None, (100.0, 80.0, 28.0)])
def test_verlet_list_bitwise(kind, box):
    sp = _verlet(box=box)
    print(f"vlist {kind} box={sp.box}", flush=True)
    tab = oracle.build_table(sp.pot)
    eng = engine(sp, tab, "mixed")
    eng.nlist()
    _check_list(eng, sp, kind)
    # positions moved beyond the skin (a shift of the whole system plus a little noise, so no
    # pair comes closer than the table reaches): the next pass rebuilds, and the list is that of
    # the new positions
    rng = np.random.default_rng(7)
    L = sp.box
    ro2 = sp.ro + np.array([1.5, -1.2, 0.7]) + rng.uniform(-0.3, 0.3, sp.ro.shape)
    ro2 = (ro2 + 0.5 * L) % L - 0.5 * L
    eng.set_positions(ro2)
    eng.utotal()
    assert eng.nlist_builds() == 2
    _check_list(eng, systems.SystemSpec(**{**sp.__dict__, "ro": ro2}), kind)
    eng.close()


def test_verlet_list_overflow_of_caller_width():
    sp = _verlet(n=1500, L=60.0)
    tab = oracle.build_table(sp.pot)
    eng = engine(sp, tab, "mixed")
    eng.nlist()
    nn, _ = eng.verlet_list("mc", 400)
    with pytest.raises(MolsimGpuError):
        eng.verlet_list("mc", int(nn.max()) - 1)
    eng.verlet_list("mc", int(nn.max()))  # fits exactly
    eng.close()


def test_skin_rule_update_order():
    """Particles 0 and 1 move by 1.0 and 1.05 A (skin 2): in this order the reference's loop
    sets r2max1 twice and leaves r2max2 = 0 (1.05 < 2, no rebuild); in the other order r2max2 =
    1.0 (2.05 > 2, rebuild).  The two largest displacements are the same in both."""
    sp = _verlet(n=2000, L=90.0, skin=2.0)
    tab = oracle.build_table(sp.pot)
    for order, want in (((1.0, 1.05), False), ((1.05, 1.0), True)):
        eng = engine(sp, tab, "fp64")
        eng.utotal()
        assert eng.nlist_builds() == 1
        ro2 = sp.ro.copy()
        ro2[0, 0] += order[0]
        ro2[1, 1] += order[1]
        ro2 = (ro2 + 45.0) % 90.0 - 45.0
        d = ro2 - sp.ro
        d -= 90.0 * np.rint(d / 90.0)
        assert oracle.nlist_skin(d, sp.drnlist)[0] == want
        eng.set_positions(ro2)
        got = eng.utotal()
        print(f"order {order}: builds {eng.nlist_builds()}", flush=True)
        assert eng.nlist_builds() == (2 if want else 1)
        ref = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro2, "list_mode": "cell"}), tab).utotal()
        assert np.linalg.norm(got["force"] - ref["force"]) / np.linalg.norm(ref["force"]) < 1e-9
        eng.close()


def test_md_rebuilds_match_reference_loop():
    sp = _verlet(n=1500, L=80.0, skin=0.6, seed=9)
    tab = oracle.build_table(sp.pot)
    mass = [20.0, 35.0]
    rng = np.random.default_rng(3)
    v0 = rng.normal(size=sp.ro.shape) * 3.0
    dt, nsteps, scale = 0.004, 30, 100.0
    eng = Engine(sp, precision="fp64", table=cabi_table(tab), force_scale=scale)
    eng.md_init(mass, v0)
    b0 = eng.nlist_builds()
    eng.md_run(nsteps, dt)
    x = eng.positions()
    xr, vr, rebuilds, x_build = vv_reference(sp, tab, mass, v0, dt, nsteps, scale)
    print(f"rebuild steps {rebuilds}; gpu builds {eng.nlist_builds() - b0}", flush=True)
    assert len(rebuilds) >= 2
    assert eng.nlist_builds() - b0 == len(rebuilds)
    dx = x - xr
    dx -= sp.box[0] * np.rint(dx / sp.box[0])
    assert np.abs(dx).max() < 1e-9
    # the list after the run is the one of the last rebuild's positions
    _check_list(eng, systems.SystemSpec(**{**sp.__dict__, "ro": x_build}), "md")
    eng.close()


def test_mixed_md_boundary_crossing_and_interval_switch():
    """Mixed path (fixed-point Verlet records relative to each particle's cell at the build):
    particles crossing the periodic boundary between builds follow the oracle-driven
    velocity-Verlet loop to the mixed tolerance; after a run with a fixed rebuild interval,
    returning to the skin rule builds the list afresh and the forces match the oracle."""
    sp = _verlet(n=1500, L=80.0, skin=0.6, seed=11)
    tab = oracle.build_table(sp.pot)
    mass = [20.0, 35.0]
    rng = np.random.default_rng(5)
    v0 = rng.normal(size=sp.ro.shape) * 3.0
    near = sp.ro[:, 0] > 0.5 * sp.box[0] - 1.0  # move out through the +x face
    v0[near, 0] = np.abs(v0[near, 0]) + 20.0
    dt, nsteps, scale = 0.004, 12, 100.0
    eng = Engine(sp, precision="mixed", table=cabi_table(tab), force_scale=scale)
    eng.md_init(mass, v0)
    eng.md_run(nsteps, dt)
    x = eng.positions()
    xr, _, rebuilds, _ = vv_reference(sp, tab, mass, v0, dt, nsteps, scale)
    crossed = int(np.count_nonzero(near & (x[:, 0] < 0.0)))
    dx = x - xr
    dx -= sp.box[0] * np.rint(dx / sp.box[0])
    print(f"{crossed} particles crossed +x; rebuild steps {rebuilds}; max |dx| {np.abs(dx).max():.2e}", flush=True)
    assert crossed > 0
    assert np.abs(dx).max() < 1e-5
    eng.nlist_interval(4)
    eng.md_run(6, dt)
    eng.nlist_interval(0)
    b = eng.nlist_builds()
    got = eng.utotal()
    assert eng.nlist_builds() == b + 1
    osp = systems.SystemSpec(**{**sp.__dict__, "ro": eng.positions(), "list_mode": "cell"})
    ref = oracle_system(osp, tab).utotal()
    ref["force"] = ref["force"] * scale
    from gpu_helpers import compare
    compare(got, ref, 1e-5, 1e-5)
    eng.close()


def test_md_fixed_interval_and_macrostep():
    sp = _verlet(n=1500, L=80.0, skin=0.6, seed=10)
    tab = oracle.build_table(sp.pot)
    eng = Engine(sp, precision="mixed", table=cabi_table(tab), force_scale=100.0)
    rng = np.random.default_rng(4)
    eng.md_init([20.0, 35.0], rng.normal(size=sp.ro.shape))
    eng.nlist_interval(5)
    b0 = eng.nlist_builds()
    eng.md_run(12, 0.002)            # steps 1..12 of the macrostep: rebuilds after 5 and 10
    assert eng.nlist_builds() - b0 == 2
    eng.md_run(3, 0.002)             # steps 13..15: after 15
    assert eng.nlist_builds() - b0 == 3
    eng.nlist_macrostep()            # the step count starts again
    eng.md_run(4, 0.002)             # steps 1..4: none
    assert eng.nlist_builds() - b0 == 3
    eng.md_run(1, 0.002)             # step 5
    assert eng.nlist_builds() - b0 == 4
    eng.close()


def test_md_overflow_during_run_is_reported():
    """A list width that fits the initial positions, then a compressed box: the first pass of
    the run overflows its list, and the run reports the overflow at its end (the flags of every
    step are kept, not only the last step's)."""
    sp = _verlet(n=1500, L=80.0, skin=1.0, seed=11)
    tab = oracle.build_table(sp.pot)
    eng = Engine(sp, precision="mixed", table=cabi_table(tab), force_scale=100.0)
    eng.nlist()
    nn, _ = eng.verlet_list("mc", 400)
    eng.set_nlist_width(int(nn.max()) + 2)
    rng = np.random.default_rng(5)
    eng.md_init([20.0, 35.0], rng.normal(size=sp.ro.shape))
    eng.md_run(3, 0.001)                         # fits
    ro = eng.positions()
    eng.set_box([72.0, 72.0, 72.0])              # density x 1.37
    eng.set_positions(ro * 0.9)
    with pytest.raises(MolsimGpuError) as ei:
        eng.md_run(6, 0.001)
    print(f"raised: {ei.value}", flush=True)
    assert ei.value.code == -11                  # MSG_ERR_NLIST
    eng.close()
