"""The Fortran drop-in: MolsimGpu module -> C-ABI -> HIP, checked against the oracle."""
import os

import numpy as np
import pytest

import oracle
from molsim_amd import systems
from fortran_io import DRIVER, parse, run, write_input
from gpu_helpers import oracle_system

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_matches_oracle(tmp_path):
    sp = systems.small_spme_system(n=400, L=48.0, nmesh=32, alpha=0.3, rcut=12.0)
    rng = np.random.default_rng(9)
    trials = []
    for _ in range(6):
        ip = int(rng.integers(sp.n))
        trials.append((ip, (sp.ro[ip] + rng.uniform(-1, 1, 3) + 24.0) % 48.0 - 24.0))
    path = tmp_path / "in.txt"
    write_input(path, sp, trials)
    print("fortran driver n=400 K=32", flush=True)
    r = run(str(path))
    assert r.returncode == 0, r.stdout + r.stderr
    got = parse(r.stdout)
    tab = oracle.build_table(sp.pot)
    osys = oracle_system(sp, tab)
    ref = osys.utotal()
    ferr = np.linalg.norm(got["force"] - ref["force"]) / np.linalg.norm(ref["force"])
    assert ferr < 1e-5
    assert abs(got["u_tot"] - ref["u_tot"]) < 1e-5 * abs(ref["u_tot"])
    for (ovl, du), (ip, rn) in zip(got["dutotal"], trials):
        d = osys.dutotal(ip, rn)
        assert bool(ovl) == d["overlap"]
        if not ovl:
            assert abs(du - d["du_tot"]) < 2e-5 + 1e-6 * abs(d["du_tot"])
    # the batched trials equal the single ones, and the cell list is MOLSIM's (grid and the
    # 1-based members in linked-list order, checksummed as sum(k * cell_ip[k]))
    assert got["dutotal_batch"] == got["dutotal"]
    onc, ocs, _ = oracle.cell_init(sp.box, sp.rcut + sp.drnlist, sp.rcut)
    head, nxt, npart, _ = oracle.cell_build(sp.ro, sp.box, onc, ocs)
    order = [ip + 1 for m in oracle.cell_members(head, nxt) for ip in m]
    assert got["cells"][0] == list(onc)
    assert got["cells"][1] == sum((k + 1) * v for k, v in enumerate(order))
    ro2 = sp.ro.copy()
    ro2[trials[0][0]] = trials[0][1]
    ref2 = oracle_system(systems.SystemSpec(**{**sp.__dict__, "ro": ro2}), tab).utotal()
    assert abs(got["u_tot_after"] - ref2["u_tot"]) < 1e-5 * abs(ref2["u_tot"])
    # the Fortran binding of MCPass against the same pass through the Python mirror
    from gpu_helpers import cabi_table
    from molsim_amd.engine import Engine, Random
    eng = Engine(sp, precision="mixed", table=cabi_table(tab))
    eng.mc_update(trials[0][0], trials[0][1])
    rng = Random(-4711)
    dsum, nacc, _ = eng.mc_pass(50, 2.0, 1e3 / (8.314472 * 298.0), rng)
    assert got["mc_pass"][0] == nacc
    assert abs(got["mc_pass"][1] - dsum) < 1e-6 * (1.0 + abs(dsum))
    assert got["rng_state"] == [rng.iseed.value, int(rng.state[0]), int(rng.state[1])]
    eng.close()
