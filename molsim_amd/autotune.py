"""Pair-kernel tile autotuner, per system-size bucket.

Two objectives: ``tune`` (latency of one system's pair kernel, for systems served alone)
and ``tune_batched`` (duration of the batched pass over all of a bucket's systems, for the
ensemble).

The pair kernel of a cell-list system takes its tile as a run-time choice
(``msg_set_tile``): ``pair_lanes`` lanes of a wavefront share one home particle's
27-cell candidate list (256 / pair_lanes homes per workgroup; 4 to 64), 0 for the
wave-per-cell-chunk kernel, or 108 / 116 for the cell-tile kernel (8 / 16 lanes per home,
the tile's neighbour cells staged in LDS; mixed precision only).  Few lanes per home suit small, sparse systems (more homes
per workgroup, fewer idle lanes); many lanes suit dense candidate lists.  The best choice
depends on the particle count and density, so the tuner times every tile on one
representative system per size bucket (the median-sized member) with per-kernel event
timing and keeps the fastest; ``apply`` then sets each system's tile from its bucket.

Results do not depend on the tile beyond the FP32 summation order of the pair sums.
"""
from __future__ import annotations

import bisect
import json
import os

TILES = (4, 8, 16, 32, 64, 0, 108, 116)
# bucket b holds systems with BUCKET_EDGES[b-1] <= n < BUCKET_EDGES[b]
BUCKET_EDGES = (300, 600, 1200, 2500, 5000, 10000, 20000, 50000, 100000)


def bucket(n: int) -> int:
    return bisect.bisect_right(BUCKET_EDGES, n)


def time_pair(engine, reps: int = 5) -> float:
    """Smallest pair-kernel time (ms) over ``reps`` UTotal passes, each kernel timed alone."""
    best = float("inf")
    engine.profile(True)
    try:
        engine.utotal(want_force=False)
        for _ in range(reps):
            engine.profile(True)
            engine.utotal(want_force=False)
            ms, n = engine.profile_read()["pair"]
            best = min(best, ms / n)
    finally:
        engine.profile(False)
    return best


def tune(engines, tiles=TILES, reps: int = 5) -> dict:
    """Table {bucket: {"tile": t, "n": representative size, "ms": {tile: ms}}}."""
    groups: dict[int, list] = {}
    for e in engines:
        groups.setdefault(bucket(e.spec.n), []).append(e)
    table = {}
    for b, members in sorted(groups.items()):
        members = sorted(members, key=lambda e: e.spec.n)
        rep = members[len(members) // 2]
        old = rep.tile()
        ms = {}
        for t in tiles:
            rep.set_tile(t)
            ms[t] = time_pair(rep, reps)
        rep.set_tile(old)
        best = min(ms, key=ms.get)
        table[b] = dict(tile=best, n=rep.spec.n, ms=ms)
    return table


def tune_batched(engines, tiles=TILES, reps: int = 5) -> dict:
    """Like ``tune`` but for throughput inside a batched pass: per size bucket, every tile is
    set on ALL members of the bucket and the bucket's batched pass (one graph launch) is
    timed; the tile with the shortest pass wins.  Table entries carry the pass time (ms)
    per tile."""
    import time

    from .engine import Batch

    groups: dict[int, list] = {}
    for e in engines:
        groups.setdefault(bucket(e.spec.n), []).append(e)
    table = {}
    for b, members in sorted(groups.items()):
        old = [e.tile() for e in members]
        ms = {}
        for t in tiles:
            for e in members:
                e.set_tile(t)
            batch = Batch(members)
            batch.utotal_async()
            batch.sync()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                batch.utotal_async()
                batch.sync()
                best = min(best, time.perf_counter() - t0)
            batch.close()
            ms[t] = 1e3 * best
        for e, t in zip(members, old):
            e.set_tile(t)
        best_t = min(ms, key=ms.get)
        table[b] = dict(tile=best_t, n=int(sorted(e.spec.n for e in members)[len(members) // 2]),
                        ms=ms, members=len(members))
    return table


def apply(engines, table: dict, default: int = 32):
    """Set every system's pair tile from its size bucket (``default`` for untuned buckets)."""
    for e in engines:
        ent = table.get(bucket(e.spec.n))
        e.set_tile(ent["tile"] if ent else default)


def save(table: dict, path: str):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1, default=str)


def load(path: str) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        raw = json.load(f)
    return {int(k): dict(v, tile=int(v["tile"])) for k, v in raw.items()}
