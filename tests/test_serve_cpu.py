"""Host logic of the request loop (no GPU): admission by footprint with LRU eviction, one
request per system in flight, FIFO order, continuous refill of free units.  Engines and
execution units are stand-ins that record what the scheduler asks of them."""
import types

import numpy as np

from molsim_amd.serve import Request, Server, request_stream, summarize


class FakeEngine:
    live = 0

    def __init__(self, sp):
        self.spec = sp
        self.closed = False
        FakeEngine.live += 1

    def close(self):
        self.closed = True
        FakeEngine.live -= 1

    def set_tile(self, t):
        self.tile = t


class FakeUnit:
    """Completes after `polls` queries; results echo the system sizes."""
    log = []

    def __init__(self, polls=2):
        self.polls = polls
        self.left = 0
        self.shape = None

    def launch(self, ef, de, trials, want_forces=True):
        assert self.left == 0
        ids = [e.spec.sid for e in ef] + [e.spec.sid for e in de]
        assert len(ids) == len(set(ids)), "a system twice in one unit"
        assert not any(e.closed for e in list(ef) + list(de))
        FakeUnit.log.append(ids)
        self.shape = (list(ef), list(de), [len(t[0]) for t in trials])
        self.left = self.polls

    def done(self):
        if self.left:
            self.left -= 1
        return self.left == 0

    def results(self):
        ef, de, nt = self.shape
        return ([dict(u_tot=float(e.spec.n)) for e in ef],
                [dict(du_tot=np.full(k, float(e.spec.n))) for e, k in zip(de, nt)])

    def close(self):
        pass


def _specs(sizes):
    out = {}
    for i, n in enumerate(sizes):
        sp = types.SimpleNamespace(n=n, sid=i, box=np.array([10.0] * 3), ro=np.zeros((n, 3)))
        out[i] = sp
    return out


def _server(specs, budget, units=2, per_unit=4):
    FakeUnit.log = []
    FakeEngine.live = 0
    return Server(specs, table=None, hbm_budget=budget, units=units, max_per_unit=per_unit,
                  unit_factory=FakeUnit, engine_factory=FakeEngine,
                  footprint_fn=lambda sp: sp.n)


def _drain(srv, reqs):
    for r in reqs:
        srv.submit(r)
    for _ in range(10000):
        srv.step()
        if srv.idle():
            return
    raise AssertionError("loop did not drain")


def test_every_request_served_once_with_its_result():
    specs = _specs([10, 20, 30, 40, 50])
    srv = _server(specs, budget=1000)
    reqs = [Request(i, i % 5, "ef" if i % 3 else "de", 0.0,
                    ips=np.arange(4), rnew=np.zeros((4, 3))) for i in range(40)]
    _drain(srv, reqs)
    assert sorted(r.rid for r in srv.done) == list(range(40))
    for r in srv.done:
        if r.kind == "ef":
            assert r.result["u_tot"] == specs[r.sys_id].n
        else:
            assert np.all(r.result["du_tot"] == specs[r.sys_id].n)
    # a system's requests complete in arrival order
    for sid in range(5):
        order = [r.rid for r in srv.done if r.sys_id == sid]
        assert order == sorted(order)


def test_units_refill_while_others_run():
    specs = _specs([10] * 8)
    srv = _server(specs, budget=1000, units=2, per_unit=3)
    _drain(srv, [Request(i, i, "ef", 0.0) for i in range(8)])
    assert [len(x) for x in FakeUnit.log] == [3, 3, 2]


def test_admission_by_footprint_and_lru_eviction():
    specs = _specs([40, 40, 40, 40])
    srv = _server(specs, budget=100, units=1, per_unit=4)
    _drain(srv, [Request(0, 0, "ef", 0.0), Request(1, 1, "ef", 0.0), Request(2, 2, "ef", 0.0)])
    # the third system did not fit beside two others in flight: it ran in a later unit
    assert FakeUnit.log[0] == [0, 1] and FakeUnit.log[1] == [2]
    assert srv.bytes <= 100 and srv.evictions >= 1
    assert srv.deferred >= 1
    assert FakeEngine.live == len(srv.resident)
    # the least recently used idle system is the one evicted
    _drain(srv, [Request(3, 3, "ef", 0.0)])
    assert 3 in srv.resident and srv.bytes <= 100


def test_oversized_system_is_an_error():
    specs = _specs([500])
    srv = _server(specs, budget=100)
    srv.submit(Request(0, 0, "ef", 0.0))
    try:
        srv.step()
    except ValueError:
        return
    raise AssertionError("expected ValueError")


def test_request_stream_poisson():
    specs = {i: types.SimpleNamespace(n=100, box=np.array([50.0] * 3), ro=np.zeros((100, 3)))
             for i in range(4)}
    reqs = request_stream(specs, qps=2000.0, duration=2.0, seed=1, frac_ef=0.5, ntrial=8)
    gaps = np.diff([r.t_arrival for r in reqs])
    assert abs(len(reqs) / 2.0 - 2000) < 150
    assert abs(gaps.mean() - 1 / 2000) < 1e-4 and abs(gaps.std() / gaps.mean() - 1) < 0.1
    de = [r for r in reqs if r.kind == "de"]
    assert 0.4 < len(de) / len(reqs) < 0.6
    assert all(r.ips.shape == (8,) and r.rnew.shape == (8, 3) for r in de)
    for r in reqs:
        r.t_done = r.t_arrival + 0.001
    s = summarize(reqs, 2.0)
    assert abs(s["latency_ms"]["p50"] - 1.0) < 1e-9
