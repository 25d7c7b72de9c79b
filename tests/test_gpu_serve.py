"""GPU tests of the request loop: execution units (segmented passes and trials over a set of
resident systems) and the continuous-batching server.

The bar: every result served in a mix -- batched with other systems, with other units in
flight, after evictions and re-admission -- equals the result of the same request served
alone by a fresh context, bit for bit (all kernels are deterministic), and the served
energies and forces match the FP64 oracle at the path's tolerance (relative force error
<= 1e-5).
"""
import numpy as np
import pytest

import oracle
from molsim_amd import serve, systems
from molsim_amd.engine import Unit
from gpu_helpers import cabi_table, compare, engine, oracle_system

pytestmark = pytest.mark.gpu

SIZES = [100, 240, 700, 1500, 3000, 5200]


@pytest.fixture(scope="module")
def pool():
    return {i: systems.served_system(40 + i, n) for i, n in enumerate(SIZES)}


@pytest.fixture(scope="module")
def table(pool):
    return oracle.build_table(pool[0].pot)


def _trials(sp, n, seed):
    rng = np.random.default_rng(seed)
    ips = rng.integers(0, sp.n, size=n)
    L = sp.box[0]
    rn = sp.ro[ips] + rng.uniform(-1.0, 1.0, size=(n, 3))
    return ips, (rn + L / 2) % L - L / 2


def _alone_ef(sp, table):
    e = engine(sp, table)
    r = e.utotal()
    e.close()
    return r


def _alone_de(sp, table, ips, rn):
    e = engine(sp, table)
    r = e.dutotal_batch(ips, rn)
    e.close()
    return r


def _same_ef(got, ref):
    return (got["u_tot"] == ref["u_tot"] and got["u_rec"] == ref["u_rec"]
            and got["virial"] == ref["virial"] and np.array_equal(got["force"], ref["force"]))


def test_unit_matches_standalone_and_oracle(pool, table):
    engs = {i: engine(sp, table) for i, sp in pool.items()}
    u1, u2 = Unit(), Unit()
    tr = {i: _trials(pool[i], 64, 70 + i) for i in (1, 3, 4)}
    print("unit 1: ef 0,1,2  de 3,4,1; unit 2 in flight beside it: ef 5", flush=True)
    u1.launch([engs[0], engs[1], engs[2]], [engs[3], engs[4], engs[1]], [tr[3], tr[4], tr[1]])
    u2.launch([engs[5]], [], [])
    ef, de = u1.results()
    ef2, _ = u2.results()
    for k, i in enumerate((0, 1, 2)):
        assert _same_ef(ef[k], _alone_ef(pool[i], table)), pool[i].name
    assert _same_ef(ef2[0], _alone_ef(pool[5], table))
    for k, i in enumerate((3, 4, 1)):
        alone = _alone_de(pool[i], table, *tr[i])
        assert np.array_equal(de[k]["overlap"], alone["overlap"])
        assert np.array_equal(de[k]["du_tot"], alone["du_tot"]), pool[i].name
    # the unit's results against the FP64 oracle
    for k, i in enumerate((0, 1)):
        ref = oracle_system(pool[i], table).utotal()
        compare(dict(ef[k], u_twob=ref["u_twob"]), ref, 1e-5, 1e-5)
    osys = oracle_system(pool[1], table)
    ips, rn = tr[1]
    for t in range(16):
        r = osys.dutotal(int(ips[t]), rn[t])
        assert de[2]["overlap"][t] == r["overlap"]
        if not r["overlap"]:
            assert abs(de[2]["du_tot"][t] - r["du_tot"]) <= 2e-5 + 1e-6 * abs(r["du_tot"])
    # the same unit again with a smaller set
    u1.launch([engs[2]], [engs[0]], [_trials(pool[0], 8, 5)])
    ef, de = u1.results()
    assert _same_ef(ef[0], _alone_ef(pool[2], table))
    for e in engs.values():
        e.close()
    u1.close()
    u2.close()


@pytest.mark.parametrize("budget_systems", [None, 2])
def test_server_mix_equals_alone(pool, table, budget_systems):
    """A Poisson stream over the pool, two units in flight; with a budget of two systems'
    footprints the loop must evict and re-admit, and results must not change."""
    ctab = cabi_table(table)
    fps = sorted(serve.footprint(sp) for sp in pool.values())
    budget = 1 << 34 if budget_systems is None else fps[-1] + fps[-2]
    srv = serve.Server(pool, ctab, hbm_budget=budget, units=2, max_per_unit=8)
    reqs = serve.request_stream(pool, qps=20000.0, duration=0.004, seed=3, frac_ef=0.5, ntrial=32)
    print(f"server: {len(reqs)} requests, budget {budget}", flush=True)
    stats = srv.run(reqs)
    assert stats["requests"] == len(reqs) and len(srv.done) == len(reqs)
    if budget_systems is not None:
        assert srv.evictions > 0 and srv.bytes <= budget
    cache = {}
    for r in srv.done:
        sp = pool[r.sys_id]
        if r.kind == "ef":
            if r.sys_id not in cache:
                cache[r.sys_id] = _alone_ef(sp, table)
            assert _same_ef(r.result, cache[r.sys_id]), (r.rid, sp.name)
        else:
            alone = _alone_de(sp, table, r.ips, r.rnew)
            assert np.array_equal(r.result["du_tot"], alone["du_tot"]), (r.rid, sp.name)
    srv.close()


def test_footprint_bounds_device_bytes(pool, table):
    """The admission estimate is an upper bound of what a context really holds after UTotal
    and DUTotal (so the budget is never exceeded)."""
    for sp in pool.values():
        e = engine(sp, table)
        e.utotal()
        e.dutotal_batch(*_trials(sp, 300, 1))
        held = e.device_bytes()
        est = serve.footprint(sp)
        print(f"{sp.name}: held {held} estimate {est}", flush=True)
        assert held <= est < 2 * held + (4 << 20)
        e.close()


def test_server_large_systems():
    """The upper end of the served range: 100 000-, 30 000- and 3 000-particle systems (SPME
    256^3, 128^3, 64^3) served in a short Poisson stream of both request kinds over two units.
    Every result equals the same request served alone, bit for bit; the 100 000-particle
    energies and forces match the FP64 oracle; the admission estimate bounds what each context
    holds."""
    big = {0: systems.served_system(71, 100000), 1: systems.served_system(72, 30000),
           2: systems.served_system(73, 3000)}
    table = oracle.build_table(big[0].pot)
    print("large pool: " + ", ".join(f"{sp.n} (K={sp.nmesh})" for sp in big.values()), flush=True)
    srv = serve.Server(big, cabi_table(table), hbm_budget=1 << 36, units=2, max_per_unit=8)
    reqs = serve.request_stream(big, qps=3000.0, duration=0.008, seed=4, frac_ef=0.5, ntrial=64)
    kinds = {(r.sys_id, r.kind) for r in reqs}
    stats = srv.run(reqs)
    print(f"served {stats['requests']} requests, p99 {stats['latency_ms']['p99']:.2f} ms", flush=True)
    assert stats["requests"] == len(reqs) and len(srv.done) == len(reqs)
    assert (0, "ef") in kinds or (0, "de") in kinds
    for sid, e in srv.resident.items():
        held, est = e.device_bytes(), serve.footprint(big[sid])
        print(f"system {sid}: held {held} estimate {est}", flush=True)
        assert held <= est < 2 * held + (4 << 20)
    cache = {}
    for r in srv.done:
        sp = big[r.sys_id]
        if r.kind == "ef":
            if r.sys_id not in cache:
                cache[r.sys_id] = _alone_ef(sp, table)
            assert _same_ef(r.result, cache[r.sys_id]), (r.rid, sp.name)
        else:
            alone = _alone_de(sp, table, r.ips, r.rnew)
            assert np.array_equal(r.result["du_tot"], alone["du_tot"]), (r.rid, sp.name)
    srv.close()
    ef0 = cache[0] if 0 in cache else _alone_ef(big[0], table)
    ref = oracle_system(big[0], table).utotal()
    compare(dict(ef0, u_twob=ref["u_twob"]), ref, 1e-5, 1e-5)
