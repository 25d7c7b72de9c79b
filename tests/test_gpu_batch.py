"""GPU tests of the batched ensemble: several resident systems of uneven size whose UTotal
passes run as parallel branches of one captured graph, and batched MC trials.

Every kernel of the path is deterministic (fixed-point mesh spreading, fixed-order
reductions), so a system's result in the batch must equal its standalone result bit for
bit; the standalone results are in turn checked against the FP64 oracle at the path's
tolerance (relative force error <= 1e-5, energies <= 1e-5 of the energy scale).
"""
import numpy as np
import pytest

import oracle
from molsim_amd import autotune, systems
from molsim_amd.engine import Batch
from gpu_helpers import compare, engine, oracle_system

pytestmark = pytest.mark.gpu

MEMBERS = [  # (seed, n, kind, f2): sizes across the mesh buckets 16^3, 32^3, 64^3
    (11, 204, "electrolyte", 0.3), (12, 260, "polyelectrolyte", 0.5),
    (13, 700, "electrolyte", 0.0), (14, 1500, "polyelectrolyte", 0.2),
    (15, 2600, "electrolyte", 0.45), (16, 4100, "polyelectrolyte", 0.1),
    (17, 9000, "electrolyte", 0.25),
]


@pytest.fixture(scope="module")
def members():
    return [systems.ensemble_member(*m) for m in MEMBERS]


@pytest.fixture(scope="module")
def table(members):
    return oracle.build_table(members[0].pot)


def _same(a, b):
    return (np.array_equal(a["force"], b["force"]) and a["u_tot"] == b["u_tot"]
            and a["u_rec"] == b["u_rec"] and np.array_equal(a["u_twob"], b["u_twob"])
            and a["virial"] == b["virial"])


def test_batch_equals_standalone(members, table):
    tiles = [8, 16, 32, 64, 0, 32, 16]
    engs = [engine(sp, table) for sp in members]
    alone = []
    for e, t in zip(engs, tiles):
        e.set_tile(t)
        assert e.tile() == t
        alone.append(e.utotal())
    b = Batch(engs)
    for rep in range(3):
        print(f"batch of {len(engs)} systems, launch {rep}", flush=True)
        got = b.utotal()
        for i, (g, a) in enumerate(zip(got, alone)):
            assert _same(g, a), (rep, members[i].name)
        en = b.energies()
        assert np.array_equal(en["u_tot"], [a["u_tot"] for a in alone])
        assert np.array_equal(en["virial"], [a["virial"] for a in alone])
    # three members against the oracle
    for i in (0, 1, 3):
        ref = oracle_system(members[i], table).utotal()
        compare(alone[i], ref, 1e-5, 1e-5)
    b.close()
    for e in engs:
        e.close()


def test_tiles_agree(members, table):
    sp = members[4]
    e = engine(sp, table)
    ref = oracle_system(sp, table).utotal()
    base = None
    for t in autotune.TILES:
        e.set_tile(t)
        r = e.utotal()
        print(f"tile {t}: n={sp.n}", flush=True)
        compare(r, ref, 1e-5, 1e-5)
        if base is None:
            base = r
        else:
            ferr = np.linalg.norm(r["force"] - base["force"]) / np.linalg.norm(base["force"])
            # the tiles differ in FP32 summation order; the cell-tile kernel (>= 100) also in
            # its coordinates (relative to the tile's first cell, up to 14 cells away), which
            # rounds the displacements differently
            assert ferr <= (5e-6 if t >= 100 else 1e-6), (t, ferr)
            assert abs(r["u_tot"] - base["u_tot"]) <= 1e-7 * abs(base["u_tot"]) + 1e-6
    e.close()


def test_autotune_table(members, table):
    engs = [engine(sp, table) for sp in members[:4]]
    tab = autotune.tune(engs, reps=2)
    assert set(tab) == {autotune.bucket(sp.n) for sp in members[:4]}
    for ent in tab.values():
        assert ent["tile"] in autotune.TILES and ent["ms"][ent["tile"]] > 0
    autotune.apply(engs, tab)
    for e in engs:
        assert e.tile() == tab[autotune.bucket(e.spec.n)]["tile"]
    for e in engs:
        e.close()


def test_autotune_batched(members, table):
    engs = [engine(sp, table) for sp in members[:5]]
    tab = autotune.tune_batched(engs, reps=2)
    assert set(tab) == {autotune.bucket(sp.n) for sp in members[:5]}
    for ent in tab.values():
        assert ent["tile"] in autotune.TILES and min(ent["ms"].values()) == ent["ms"][ent["tile"]]
    autotune.apply(engs, tab)
    b = Batch(engs)
    got = b.utotal()
    for e, g in zip(engs, got):  # tuned batch still equals the standalone result
        alone = e.utotal()
        assert _same(g, alone), e.spec.name
    b.close()
    for e in engs:
        e.close()


def _trials(sp, n, seed):
    rng = np.random.default_rng(seed)
    ips = rng.integers(0, sp.n, size=n)
    L = sp.box[0]
    rn = sp.ro[ips] + rng.uniform(-1.0, 1.0, size=(n, 3))
    return ips, (rn + L / 2) % L - L / 2


def test_batch_dutotal_and_commits(members, table):
    sub = members[:5]
    engs = [engine(sp, table) for sp in sub]
    b = Batch(engs)
    b.utotal()
    trials = [_trials(sp, 48, 100 + i) for i, sp in enumerate(sub)]
    got = b.dutotal([t[0] for t in trials], [t[1] for t in trials])
    for i, (e, (ips, rn)) in enumerate(zip(engs, trials)):
        alone = e.dutotal_batch(ips, rn)
        assert np.array_equal(got[i]["overlap"], alone["overlap"])
        assert np.array_equal(got[i]["du_tot"], alone["du_tot"]), sub[i].name
    # the smallest member's trials against the oracle
    osys = oracle_system(sub[0], table)
    ips, rn = trials[0]
    nchk = 0
    for t in range(len(ips)):
        ref = osys.dutotal(int(ips[t]), rn[t])
        assert got[0]["overlap"][t] == ref["overlap"]
        if not ref["overlap"]:
            assert abs(got[0]["du_tot"][t] - ref["du_tot"]) <= 2e-5 + 1e-6 * abs(ref["du_tot"])
            nchk += 1
    assert nchk >= 24
    # accept one non-overlapping trial in two systems, then batch again: equal to fresh systems
    ros = [sp.ro.copy() for sp in sub]
    for i in (0, 2):
        t = int(np.nonzero(~got[i]["overlap"])[0][0])
        ip, r = int(trials[i][0][t]), trials[i][1][t]
        engs[i].mc_update(ip, r)
        ros[i][ip] = r
    res = b.utotal()
    for i in (0, 2):
        fresh = engine(systems.SystemSpec(**{**sub[i].__dict__, "ro": ros[i]}), table)
        assert _same(res[i], fresh.utotal()), sub[i].name
        fresh.close()
    b.close()
    for e in engs:
        e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_mc_pass_matches_single_chains(table, precision):
    """One MCPass per system with the chains side by side (one workgroup each, one launch per
    SPME order; batched UTotal passes recompute phi between launches): every system makes the
    decisions its chain makes alone, with the same generator sequence.  FP64: the same
    decisions, final positions and generator state; mixed: at least 99 % of the decisions.
    One member runs at order 5, so two kernel instances share each launch round."""
    from molsim_amd.engine import Random
    sub = [systems.ensemble_member(11, 204, "electrolyte", 0.3),
           systems.ensemble_member(12, 260, "polyelectrolyte", 0.5),
           systems.ensemble_member(13, 700, "electrolyte", 0.0, order=5),
           systems.ensemble_member(14, 1500, "polyelectrolyte", 0.2),
           systems.ensemble_member(15, 2600, "electrolyte", 0.45)]
    ntrial, dtran, beta = 200, 1.5, systems.beta_kjmol()
    seeds = [-101 - k for k in range(len(sub))]
    alone = []
    for sp, seed in zip(sub, seeds):
        e = engine(sp, table, precision)
        g = Random(seed)
        _, na, ev = e.mc_pass(ntrial, dtran, beta, g)
        alone.append((na, ev, e.positions(), g.get_state()))
        e.close()
    print(f"batched MC pass {precision}: {len(sub)} systems x {ntrial} trials", flush=True)
    engs = [engine(sp, table, precision) for sp in sub]
    b = Batch(engs)
    rngs = [Random(s) for s in seeds]
    _, nacc, ev = b.mc_pass(ntrial, dtran, beta, rngs)
    for k, (na, ev1, ro1, st1) in enumerate(alone):
        agree = float(np.mean(ev[k] == ev1))
        print(f"  system {k} (n={sub[k].n}): accepted {nacc[k]} (alone {na}), agreement {agree:.3f}", flush=True)
        assert 0 < nacc[k] < ntrial
        if precision == "fp64":
            assert np.array_equal(ev[k], ev1)
            assert nacc[k] == na
            assert np.abs(engs[k].positions() - ro1).max() < 1e-12
            assert rngs[k].get_state() == st1
        else:
            assert agree >= 0.99
    # the systems' state after the pass is that of their final positions
    got = b.utotal()
    for k in (0, 2):
        ref = oracle_system(systems.SystemSpec(**{**sub[k].__dict__, "ro": engs[k].positions()}), table).utotal()
        tol = 1e-9 if precision == "fp64" else 1e-5
        compare(got[k], ref, tol, tol)
    b.close()
    for e in engs:
        e.close()
