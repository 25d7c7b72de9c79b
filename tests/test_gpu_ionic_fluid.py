"""The 1 048 576-ion fluid at full size (SPME 256^3 order 6, Verlet list rc 18 A + skin 2 A):
UTotal of the mixed and FP64 paths against the oracle, and the Verlet list bit for bit against
the oracle's SetVListMD rows.  At this size the cell grid (40^3 cells) and the mesh (16384
bricks) are too large for the fused one-launch sort, so the separate scan + scatter path runs;
so do the 256^3 FFT instantiations.  A smaller system on the same mesh checks the separate
brick path with a cell list and the captured graph.

The oracle UTotal at this size runs in its OpenMP build (the same code with its loops shared
out over the host's cores; tests/test_oracle_numerics.py checks both builds agree)."""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu


def _oracle(sp, tab, list_mode="cell"):
    return oracle_system(systems.SystemSpec(**{**sp.__dict__, "list_mode": list_mode}), tab)


def test_ionic_fluid_full_size():
    sp = systems.ionic_fluid()
    print(f"ionic fluid n={sp.n} K={sp.nmesh} box={sp.box[0]:.2f}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = _oracle(sp, tab).utotal(omp=True)
    print("oracle done", flush=True)
    for precision, tol in (("mixed", 1e-5), ("fp64", 1e-9)):
        eng = engine(sp, tab, precision)
        compare(eng.utotal(), ref, tol, tol)
        if precision == "mixed":
            nn, jl = eng.verlet_list("md", 128)
            onn, ojl = oracle.vlist("md", sp.ro, sp.box, sp.rcut, sp.drnlist, maxn=128, cells=True, table=True)
            print(f"md list: {int(nn.sum())} pairs, longest row {int(nn.max())}", flush=True)
            assert np.array_equal(nn, onn)
            assert np.array_equal(jl, ojl)
            del jl, ojl
        eng.close()


def test_dense_mesh_separate_brick_sort():
    """K = 256 over 20 000 ions: 16 384 mesh bricks (separate brick scan and scatter) with a
    cell list (fused cell sort) in the captured graph."""
    sp = systems.small_spme_system(seed=12, n=20000, L=400.0, nmesh=256, alpha=0.15, rcut=20.0)
    print(f"dense mesh n={sp.n} K={sp.nmesh}", flush=True)
    tab = oracle.build_table(sp.pot)
    ref = _oracle(sp, tab).utotal(omp=True)
    eng = engine(sp, tab, "mixed")
    compare(eng.utotal(), ref, 1e-5, 1e-5)
    compare(eng.utotal(), ref, 1e-5, 1e-5)   # a graph replay
    eng.close()
