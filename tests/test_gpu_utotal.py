"""GPU parity of UTotal (cell list + tabulated pair loop + SPME) against the FP64 oracle.

Tolerances: the mixed path (FP32 pair/spread/FFT arithmetic, FP64 accumulation) must give
forces within 1e-5 relative 2-norm error of the oracle and energies within 1e-5 of the
energy scale; the plain FP64 path within 1e-9.  Cell lists must match bit for bit.
"""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import Engine, MolsimGpuError
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu

MIXED_FTOL = 1e-5
MIXED_ETOL = 1e-5
FP64_TOL = 1e-9


def _small(**kw):
    args = dict(n=400, L=48.0, nmesh=32, alpha=0.3, rcut=12.0)
    args.update(kw)
    return systems.small_spme_system(**args)


@pytest.mark.parametrize("precision,ftol,etol", [("mixed", MIXED_FTOL, MIXED_ETOL),
                                                 ("fp64", FP64_TOL, FP64_TOL)])
@pytest.mark.parametrize("case", ["cell_spme", "cell_spme_nosurf", "all_spme", "cell_k16",
                                  "cell_order4", "three_types", "dense_bricks", "big_charges",
                                  "clustered"])
def test_utotal_small(case, precision, ftol, etol):
    """dense_bricks: about 156 particles per 32x8x8 mesh brick (about 1250 spread candidates per
    brick: several candidate rounds per wavefront, passes of more than 32 contributors, and a
    coarser 32-bit fixed-point scale of the FP32 spread).  big_charges: charges +-20 (the scale moves
    down by 2^5).  clustered: 250 ions in a 25 A cube inside a 60 A box, a few bricks holding
    all of them."""
    kw = {}
    if case == "cell_spme_nosurf":
        kw = dict(lsurf=False)
    elif case == "all_spme":
        kw = dict(list_mode="all", n=200, L=30.0, rcut=14.0)
    elif case == "cell_k16":
        kw = dict(nmesh=16, alpha=0.2, n=150)
    elif case == "cell_order4":
        kw = dict(order=4, nmesh=64)
    elif case == "three_types":
        kw = dict(charges=(1.0, -2.0, 1.0), n=399)
    elif case == "dense_bricks":
        kw = dict(n=2500, L=60.0)
    elif case == "big_charges":
        kw = dict(charges=(20.0, -20.0))
    elif case == "clustered":
        kw = dict(n=250, L=25.0)
    sp = _small(**kw)
    if case == "clustered":  # the ions as inserted into 25 A, now in a 60 A box (mesh 32^3)
        sp = systems.SystemSpec(**{**sp.__dict__, "box": np.array([60.0, 60.0, 60.0])})
    if case == "dense_bricks":
        assert sp.n > 48 * (sp.nmesh // 16) * (sp.nmesh // 8) ** 2
    print(f"case={case} precision={precision} n={sp.n} K={sp.nmesh} order={sp.order}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    gpu = eng.utotal()
    compare(gpu, ref, ftol, etol)
    eng.close()


def test_cell_list_bitwise():
    sp = _small(n=3000, L=100.0, rcut=9.0)
    print(f"cell list n={sp.n}", flush=True)
    eng = Engine(sp)
    eng.nlist()
    nc, start, ips = eng.cell_list()
    onc, ocs, _ = oracle.cell_init(sp.box, sp.rcut + sp.drnlist, sp.rcut)
    assert np.array_equal(nc, onc)
    head, nxt, npart, _ = oracle.cell_build(sp.ro, sp.box, onc, ocs)
    members = oracle.cell_members(head, nxt)
    assert np.array_equal(np.diff(start), npart)
    for c, m in enumerate(members):
        assert list(ips[start[c]:start[c + 1]]) == m
    eng.close()


def test_polyelectrolyte_full_size():
    """The 32768-particle polyelectrolyte, SPME 128^3 order 6, rc 20 A cell list."""
    sp = systems.polyelectrolyte()
    print(f"polyelectrolyte n={sp.n} K={sp.nmesh}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab)
    gpu = eng.utotal()
    errs = compare(gpu, ref, MIXED_FTOL, MIXED_ETOL)
    # per-particle worst case is reported, not bounded by the global metric
    fe = np.linalg.norm(gpu["force"] - ref["force"], axis=1) / np.linalg.norm(ref["force"], axis=1).mean()
    print(f"max per-particle force error / mean |F|: {fe.max():.2e}", flush=True)
    eng.close()


def test_deterministic_and_graph_replay():
    sp = _small(n=2000, L=80.0, nmesh=64)
    eng = Engine(sp)
    runs = [eng.utotal() for _ in range(4)]       # first call captures the graph, then replays
    eng.profile(True)                              # eager launches (events between kernels)
    runs.append(eng.utotal())
    eng.profile(False)
    for r in runs[1:]:
        assert np.array_equal(r["force"], runs[0]["force"])
        assert r["u_tot"] == runs[0]["u_tot"] and r["virial"] == runs[0]["virial"]
    eng.close()


def test_positions_update_between_calls():
    sp = _small(n=600, L=50.0)
    tab = oracle.build_table(sp.pot)
    eng = engine(sp, tab)
    eng.utotal()
    rng = np.random.default_rng(4)
    ro2 = sp.ro + rng.uniform(-0.3, 0.3, sp.ro.shape)
    ro2 = (ro2 + 25.0) % 50.0 - 25.0
    eng.set_positions(ro2)
    gpu = eng.utotal()
    sp2 = systems.SystemSpec(**{**sp.__dict__, "ro": ro2})
    ref = oracle_system(sp2, tab).utotal()
    compare(gpu, ref, MIXED_FTOL, MIXED_ETOL)
    eng.close()


def test_r2umin_error_is_reported():
    sp = _small(n=50, L=30.0)
    pot = dict(sp.pot, rad=[0.0, 0.0])          # no hard core: the LJ table ends at r2umin
    ro = sp.ro.copy()
    ro[1] = ro[0] + np.array([0.5, 0.0, 0.0])    # far below r2umin
    sp2 = systems.SystemSpec(**{**sp.__dict__, "ro": ro, "pot": pot})
    eng = Engine(sp2)
    with pytest.raises(MolsimGpuError) as e:
        eng.utotal()
    assert e.value.code == -2
    eng.close()


@pytest.mark.parametrize("tile", [116, 108, 4, 8, 64, 0])
@pytest.mark.parametrize("case", ["poly", "collapsed", "dense"])
def test_pair_tiles_match_oracle(case, tile):
    """Every pair-kernel tile (the cell-tile kernel with 16 or 8 lanes per home, the
    neighbour-parallel and the wave-per-cell-chunk kernels) against the oracle: the 32 768-ion
    polyelectrolyte (dense chains), a box with a dimension collapsed to one cell, and cells dense
    enough that tiles overflow their LDS staging area and read candidates from memory."""
    if case == "poly":
        sp = systems.polyelectrolyte()
    elif case == "collapsed":
        from scipy.spatial import cKDTree
        sp = _small(n=800, L=60.0, rcut=9.0)
        box = np.array([60.0, 60.0, 16.0])
        rng = np.random.default_rng(12)
        pts = rng.uniform(0.0, 1.0, (sp.n, 3)) * box
        for _ in range(200):   # reinsert until no two ions are closer than 3 A
            pairs = cKDTree(pts, boxsize=box).query_pairs(3.0, output_type="ndarray")
            if not len(pairs):
                break
            bad = np.unique(pairs[:, 1])
            pts[bad] = rng.uniform(0.0, 1.0, (len(bad), 3)) * box
        sp = systems.SystemSpec(**{**sp.__dict__, "ro": pts - 0.5 * box, "box": box})
    else:
        sp = _small(n=2200, L=56.0, rcut=18.0, alpha=0.2)   # 3x3x3 cells, 2200 > 2048 records
    print(f"tiles case={case} tile={tile} n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab)
    eng.set_tile(tile)
    compare(eng.utotal(), ref, MIXED_FTOL, MIXED_ETOL)
    compare(eng.utotal(), ref, MIXED_FTOL, MIXED_ETOL)   # graph replay
    eng.close()


@pytest.mark.parametrize("precision,ftol,etol", [("mixed", MIXED_FTOL, MIXED_ETOL), ("fp64", FP64_TOL, FP64_TOL)])
def test_crowded_brick(precision, ftol, etol):
    """700 of 800 ions in one of the four 16 x 8 x 8 mesh bricks (K = 16, L = 60 A): one brick
    holds nearly all the charge the spread deposits and a few cells hold most of the pair
    work; against the oracle, twice (graph replay)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(17)
    L = 60.0
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([60.0, 30.0, 30.0])
    pts = np.concatenate([rng.uniform(lo, hi, (700, 3)), rng.uniform(0.0, L, (100, 3))])
    pts[700:, 1] = 30.0 + (pts[700:, 1] % 30.0)   # the other 100 outside that brick
    for _ in range(300):
        pairs = cKDTree(pts % L, boxsize=L).query_pairs(3.0, output_type="ndarray")
        if not len(pairs):
            break
        bad = np.unique(pairs[:, 1])
        inside = bad < 700
        pts[bad[inside]] = rng.uniform(lo, hi, (int(inside.sum()), 3))
        out = bad[~inside]
        pts[out] = rng.uniform(0.0, L, (len(out), 3))
        pts[out, 1] = 30.0 + (pts[out, 1] % 30.0)
    assert not len(cKDTree(pts % L, boxsize=L).query_pairs(3.0, output_type="ndarray"))
    sp = _small(n=800, L=L, nmesh=16, rcut=12.0)
    sp = systems.SystemSpec(**{**sp.__dict__, "ro": (pts % L) - 0.5 * L})
    print(f"crowded brick {precision}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, ftol, etol)
    compare(eng.utotal(), ref, ftol, etol)
    eng.close()


@pytest.mark.parametrize("precision,ftol,etol", [("mixed", MIXED_FTOL, MIXED_ETOL),
                                                 ("fp64", FP64_TOL, FP64_TOL)])
def test_brick_class_boundaries(precision, ftol, etol):
    """A 13 x 13 x 14 ion lattice in a 64 A box with a 64^3 mesh (1 A spacing, order 6, bricks
    of 32 x 8 x 8 points sorted by class): the B-spline base points take every offset 0..7 from
    a brick's low side in y, and offsets on both sides of the order - 1 = 5 point limit that
    decides whether a particle of the next brick reaches back into a brick in x (2 and 5) and z
    (4 and 6).  A class boundary drawn one point too tight drops contributions; energy and
    forces against the oracle."""
    sp = _small(n=2366, L=64.0, nmesh=64)
    ax = [np.arange(13) * (64.0 / 13), np.arange(13) * (64.0 / 13), np.arange(14) * (64.0 / 14)]
    g = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    ro = g + 0.37 - 32.0
    par = np.stack(np.meshgrid(np.arange(13), np.arange(13), np.arange(14), indexing="ij"), -1).reshape(-1, 3).sum(1)
    ipt = (par % 2).astype(np.int32)
    az = np.where(ipt == 0, 1.0, -1.0)
    sp = systems.SystemSpec(**{**sp.__dict__, "ro": ro, "ipt": ipt, "az": az})
    base = np.floor(ro).astype(int) % 64  # dkdr = K / L = 1 per A
    assert set(base[:, 1] % 8) == set(range(8)) and {4, 6} <= set(base[:, 2] % 8)
    assert {2, 5} <= set(base[:, 0] % 32)
    print(f"brick classes precision={precision} n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = oracle_system(sp, tab).utotal()
    eng = engine(sp, tab, precision)
    compare(eng.utotal(), ref, ftol, etol)
