"""msg_dutotal_batch_noalloc (the entry point that allocates nothing) against msg_dutotal_batch
through the C-ABI: bit for bit on two calls in a row, the second with fewer trials, for an SPME
cell-list system and a standard-Ewald all-pairs system; sizes beyond the scratch are refused."""
import numpy as np
import pytest

import oracle
from molsim_amd import systems
from molsim_amd.engine import MolsimGpuError
from gpu_helpers import engine

pytestmark = pytest.mark.gpu


def _trials(sp, n, seed):
    rng = np.random.default_rng(seed)
    ips = rng.integers(0, sp.n, size=n)
    L = sp.box
    rn = sp.ro[ips] + rng.uniform(-2.0, 2.0, (n, 3))
    return ips, (rn + 0.5 * L) % L - 0.5 * L


@pytest.mark.parametrize("kind", ["spme_cell", "std_all"])
def test_noalloc_equals_allocating_entry(kind):
    if kind == "spme_cell":
        sp = systems.small_spme_system(seed=13, n=3000, L=100.0, nmesh=32, alpha=0.3, rcut=12.0)
    else:
        sp = systems.electrolyte_primitive(seed=14, n=600, L=70.0)
    print(f"noalloc {kind} n={sp.n}", flush=True)
    tab = oracle.build_table(sp.pot)
    for precision in ("mixed", "fp64"):
        eng = engine(sp, tab, precision)
        for n, seed in ((3000, 1), (700, 2)):   # the second call has fewer trials
            ips, rn = _trials(sp, n, seed)
            want = eng.dutotal_batch(ips, rn)
            got = eng.dutotal_batch_noalloc(ips, rn)
            assert np.array_equal(got["overlap"], want["overlap"])
            assert np.array_equal(got["du_tot"], want["du_tot"])
            assert np.array_equal(got["du_rec"], want["du_rec"])
        with pytest.raises(MolsimGpuError):
            eng.dutotal_batch_noalloc(np.zeros(16385, dtype=np.int64), np.zeros((16385, 3)))
        eng.close()


def test_noalloc_needs_spme_tables_first():
    sp = systems.small_spme_system(seed=15, n=500, L=60.0, nmesh=32, alpha=0.3, rcut=12.0)
    tab = oracle.build_table(sp.pot)
    eng = engine(sp, tab, "mixed")
    ips, rn = _trials(sp, 10, 3)
    with pytest.raises(MolsimGpuError) as ei:
        eng.dutotal_batch_noalloc(ips, rn)
    assert ei.value.code == -12  # MSG_ERR_STATE
    want = eng.dutotal_batch(ips, rn)
    got = eng.dutotal_batch_noalloc(ips, rn)
    assert np.array_equal(got["du_tot"], want["du_tot"])
    eng.close()
