"""Request loop over device-resident systems: continuous batching with HBM-footprint admission.

Requests ask for a full UTotal (energy + forces) of a system, or for the energy changes of a
batch of single-particle trial moves (DUTotal) of a system.  The loop keeps every system it
has admitted resident on the device (positions, cell lists, meshes, tables), so a request
costs only its kernels.

Scheduling (iteration-level, as in continuous batching):

* Requests wait in arrival order.  Whenever an execution unit (``Unit``: a stream pair with
  its own argument arena) is free, the scheduler fills it with every waiting request whose
  system is not already in flight, up to ``max_per_unit`` systems, and launches it; the
  unit's UTotal systems form one segmented pass (one launch per stage over all of them) and
  its trial systems one segmented trial launch.  Requests arriving while units run join the
  next free unit, so several units overlap on the device.
* A system enters the device only if its footprint fits the HBM budget: the estimate
  ``footprint(spec)`` is checked before the context is created; idle resident systems are
  evicted least-recently-used first to make room; a request whose system cannot be made to
  fit stays queued until memory frees (in-flight systems are never evicted).

Every timestamp comes from one monotonic clock; latency = completion - arrival.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine, Unit


def footprint(spec) -> int:
    """Upper estimate (bytes) of the device memory a mixed-precision cell-list SPME context
    holds after UTotal and DUTotal have run: per particle ~ 250 B (positions, charges, cell
    and brick orderings and ranks, forces, B-spline data, pair partials), per mesh point ~ 10 B (FP32
    mesh, half-complex mesh, influence function, FP64 trial kernel), per cell 40 B, plus
    fixed partial-sum blocks and tables."""
    n = spec.n
    K = max(spec.nmesh, 1)
    H = K // 2 + 1
    rc = spec.rcut + spec.drnlist
    ncell = 1
    for d in range(3):
        ncell *= max(1, int(spec.box[d] // rc))
    mesh = 4 * K ** 3 + 8 * H * K * K + 4 * H ** 3 + 8 * H ** 3 + 64 * K
    return int(1.15 * (270 * n + mesh + 40 * ncell + (2 << 20)))


@dataclass
class Request:
    rid: int
    sys_id: int
    kind: str                      # "ef" (energy + forces) or "de" (trial moves)
    t_arrival: float
    ips: np.ndarray | None = None  # de: particle indices (0-based)
    rnew: np.ndarray | None = None  # de: trial positions (n, 3)
    t_launch: float = 0.0
    t_done: float = 0.0
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class Server:
    """The request loop.  ``specs[sys_id]`` describes every system a request may target;
    ``table`` is the shared pair table (C-ABI convention)."""

    def __init__(self, specs: dict, table: dict, hbm_budget: int, units: int = 2,
                 max_per_unit: int = 64, precision: str = "mixed", tiles: dict | None = None,
                 clock=time.perf_counter, unit_factory=Unit, engine_factory=None,
                 footprint_fn=footprint):
        self.specs = specs
        self.engine_factory = engine_factory or (
            lambda sp: Engine(sp, precision=self.precision, table=self.table))
        self.footprint = footprint_fn
        self.table = table
        self.budget = int(hbm_budget)
        self.max_per_unit = max_per_unit
        self.precision = precision
        self.tiles = tiles or {}
        self.clock = clock
        self.units = [unit_factory() for _ in range(units)]
        self.inflight: dict[int, list[Request]] = {}    # unit index -> requests
        self.resident: collections.OrderedDict[int, Engine] = collections.OrderedDict()
        self.bytes = 0
        self.reserved: dict[int, int] = {}               # sys_id -> admitted footprint
        self.busy: set[int] = set()
        self.queue: collections.deque[Request] = collections.deque()
        self.done: list[Request] = []
        self.evictions = 0
        self.deferred = 0

    # -- residency ----------------------------------------------------------------------------
    def _admit_system(self, sys_id: int, protect=()) -> Engine | None:
        """The resident context of sys_id, made resident if its footprint fits (evicting idle
        systems not in ``protect``); None when it cannot fit now."""
        eng = self.resident.get(sys_id)
        if eng is not None:
            self.resident.move_to_end(sys_id)
            return eng
        need = self.footprint(self.specs[sys_id])
        if need > self.budget:
            raise ValueError(f"system {sys_id} needs {need} B, more than the whole budget")
        # evict idle systems, least recently used first
        for victim in list(self.resident):
            if self.bytes + need <= self.budget:
                break
            if victim in self.busy or victim in protect:
                continue
            self.resident.pop(victim).close()
            self.bytes -= self.reserved.pop(victim)
            self.evictions += 1
        if self.bytes + need > self.budget:
            return None
        eng = self.engine_factory(self.specs[sys_id])
        if sys_id in self.tiles:
            eng.set_tile(self.tiles[sys_id])
        self.resident[sys_id] = eng
        held = max(need, int(getattr(eng, "device_bytes", lambda: 0)()))
        self.reserved[sys_id] = held
        self.bytes += held
        return eng

    # -- system state in MOLSIM's configuration files ------------------------------------------
    def ingest_cnf(self, sys_id: int, path, **read_kw):
        """Replace a system's positions (and box) by a MOLSIM .cnf file (molsim_amd.cnf).  A
        resident system with an unchanged box is updated in place; one whose box changed is
        evicted, so its next request builds it afresh.  A system in flight cannot change."""
        from . import cnf
        if sys_id in self.busy:
            raise RuntimeError(f"system {sys_id} is in flight")
        spec = cnf.apply_to_spec(self.specs[sys_id], cnf.read_cnf(path, np_expected=self.specs[sys_id].n,
                                                                    **read_kw))
        same_box = np.array_equal(spec.box, self.specs[sys_id].box)
        self.specs[sys_id] = spec
        eng = self.resident.get(sys_id)
        if eng is not None:
            if same_box:
                eng.spec = spec
                eng.set_positions(spec.ro)
            else:
                self.resident.pop(sys_id).close()
                self.bytes -= self.reserved.pop(sys_id)

    def emit_cnf(self, sys_id: int, path, istep: int = 0, rng=None):
        """Write a system's current state as a MOLSIM .cnf file."""
        from . import cnf
        cnf.write_cnf(path, cnf.from_spec(self.specs[sys_id], istep, rng))

    # -- scheduling ---------------------------------------------------------------------------
    def submit(self, req: Request):
        self.queue.append(req)

    def _complete(self):
        for ui, reqs in list(self.inflight.items()):
            u = self.units[ui]
            if not u.done():
                continue
            ef, de = u.results()
            now = self.clock()
            ief = ide = 0
            for r in reqs:
                if r.kind == "ef":
                    r.result = ef[ief]
                    ief += 1
                else:
                    r.result = de[ide]
                    ide += 1
                r.t_done = now
                self.busy.discard(r.sys_id)
                self.done.append(r)
            del self.inflight[ui]

    def _launch(self):
        for ui, u in enumerate(self.units):
            if ui in self.inflight or not self.queue:
                continue
            picked, ef, de, trials, taken = [], [], [], [], set()
            keep = collections.deque()
            while self.queue and len(taken) < self.max_per_unit:
                r = self.queue.popleft()
                if r.sys_id in self.busy or r.sys_id in taken:
                    keep.append(r)
                    continue
                eng = self._admit_system(r.sys_id, protect=taken)
                if eng is None:             # no memory for it yet: stays queued
                    self.deferred += 1
                    keep.append(r)
                    continue
                taken.add(r.sys_id)
                picked.append(r)
                if r.kind == "ef":
                    ef.append(eng)
                else:
                    de.append(eng)
                    trials.append((r.ips, r.rnew))
            keep.extend(self.queue)
            self.queue = keep
            if not picked:
                return
            # requests are reported in the unit's result order: ef first, then de
            picked.sort(key=lambda r: r.kind != "ef")
            u.launch(ef, de, trials, want_forces=True)
            now = self.clock()
            for r in picked:
                r.t_launch = now
                self.busy.add(r.sys_id)
            self.inflight[ui] = picked

    def step(self):
        """Collect finished units, then fill free units from the queue."""
        self._complete()
        self._launch()

    def idle(self) -> bool:
        return not self.queue and not self.inflight

    def run(self, requests: list[Request], spin: float = 20e-6) -> dict:
        """Replay ``requests`` (t_arrival relative to the start) in real time."""
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        t0 = self.clock()
        for r in reqs:
            r.t_arrival += t0
        i = 0
        while i < len(reqs) or not self.idle():
            now = self.clock()
            while i < len(reqs) and reqs[i].t_arrival <= now:
                self.submit(reqs[i])
                i += 1
            self.step()
            if i < len(reqs) and self.idle() and reqs[i].t_arrival > self.clock() + spin:
                time.sleep(max(0.0, reqs[i].t_arrival - self.clock() - spin))
        t1 = self.clock()
        return summarize(reqs, t1 - t0)

    def close(self):
        for u in self.units:
            u.close()
        for e in self.resident.values():
            e.close()
        self.resident.clear()


def serve_serial(engines: dict, requests: list[Request], spin: float = 20e-6) -> dict:
    """The one-at-a-time baseline served end to end: requests replayed in real time and
    executed in arrival order, each alone on its system's resident context (energy + forces:
    one UTotal with the forces copied to the host; trials: one DUTotal call per move, as
    MOLSIM's caller makes them).  Latency runs from arrival to the result on the host."""
    reqs = sorted(requests, key=lambda r: r.t_arrival)
    t0 = time.perf_counter()
    for r in reqs:
        r.t_arrival += t0
    for r in reqs:
        wait = r.t_arrival - time.perf_counter()
        if wait > spin:
            time.sleep(wait - spin)
        while time.perf_counter() < r.t_arrival:
            pass
        e = engines[r.sys_id]
        r.t_launch = time.perf_counter()
        if r.kind == "ef":
            r.result = e.utotal()
        else:
            du = np.zeros(len(r.ips))
            for k in range(len(r.ips)):
                du[k] = e.dutotal(int(r.ips[k]), r.rnew[k])["du_tot"]
            r.result = {"du_tot": du}
        r.t_done = time.perf_counter()
    return summarize(reqs, time.perf_counter() - t0)


def summarize(reqs, wall: float) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    out = {"requests": len(reqs), "wall_s": wall, "throughput_rps": len(reqs) / wall,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "mean": float(lat.mean()), "max": float(lat.max())}}
    for kind in ("ef", "de"):
        lk = np.array([r.latency for r in reqs if r.kind == kind]) * 1e3
        if lk.size:
            out[f"latency_ms_{kind}"] = {"p50": float(np.percentile(lk, 50)),
                                         "p99": float(np.percentile(lk, 99))}
    return out


def request_stream(specs: dict, qps: float, duration: float, seed: int = 9, frac_ef: float = 0.5,
                   ntrial: int = 256, dmax: float = 1.0) -> list[Request]:
    """Poisson arrivals at rate ``qps`` for ``duration`` seconds; each request targets a system
    drawn uniformly from ``specs`` (whose sizes are log-uniform) and asks for energy + forces
    (probability frac_ef) or for ``ntrial`` single-particle trial moves of at most dmax per
    axis."""
    rng = np.random.default_rng(seed)
    ids = sorted(specs)
    t, rid, out = 0.0, 0, []
    while True:
        t += rng.exponential(1.0 / qps)
        if t > duration:
            break
        sid = ids[int(rng.integers(len(ids)))]
        if rng.random() < frac_ef:
            out.append(Request(rid, sid, "ef", t))
        else:
            sp = specs[sid]
            ips = rng.integers(0, sp.n, size=ntrial)
            L = sp.box[0]
            rn = sp.ro[ips] + rng.uniform(-dmax, dmax, size=(ntrial, 3))
            out.append(Request(rid, sid, "de", t, ips=ips, rnew=(rn + L / 2) % L - L / 2))
        rid += 1
    return out


def serial_baseline(specs: dict, table: dict, requests: list[Request], precision: str = "fp64",
                    clock=time.perf_counter) -> dict:
    """The plain one-at-a-time path: every request served alone, in arrival order, by a
    resident context of its system (no batching, no overlap): energy + forces by one UTotal
    with forces returned; trial batches one DUTotal call per trial.  Each service time is
    measured; the latencies follow from the arrival times by the single-server FIFO
    recursion start = max(arrival, previous end)."""
    engines: dict[int, Engine] = {}
    service = []
    for r in sorted(requests, key=lambda q: q.t_arrival):
        eng = engines.get(r.sys_id)
        if eng is None:
            eng = engines[r.sys_id] = Engine(specs[r.sys_id], precision=precision, table=table)
            eng.utotal()                       # residency set-up is not request time
        t0 = clock()
        if r.kind == "ef":
            eng.utotal()
        else:
            for k in range(len(r.ips)):
                eng.dutotal(int(r.ips[k]), r.rnew[k])
        service.append((r, clock() - t0))
    end, lat = 0.0, []
    for r, s in service:
        start = max(r.t_arrival, end)
        end = start + s
        lat.append(end - r.t_arrival)
    for e in engines.values():
        e.close()
    lat = np.array(lat) * 1e3
    busy = sum(s for _, s in service)
    span = max(end, max(r.t_arrival for r in requests))
    return {"requests": len(requests), "throughput_rps": len(requests) / span,
            "capacity_rps": len(requests) / busy, "utilisation": busy / span,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "mean": float(lat.mean()), "max": float(lat.max())}}
