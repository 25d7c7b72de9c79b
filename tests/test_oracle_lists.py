"""Pinning of the oracle's Verlet lists, skin test and linked cell-list update (CPU only).

* vlist (SetVListMD / SetVListMC, Src/nlist.F90:724-759,820-884): the pairs within
  rcut + drnlist found by numpy brute force with a different minimum-image formula
  (d - L * rint(d / L)); every pair exactly once in the MD half lists and in the row the parity
  rule names; MC rows are the symmetric closure; rows ascending; a pair exactly at
  rcut + drnlist is excluded (r2 < rs2cut); the cell-list form equals the O(N^2) form.
* nlist_skin (Src/nlist.F90:600-612): hand-computed r2max1 / r2max2 for sequences where the
  reference's update order gives something other than the two largest values.
* LinkedCellList (SetCellList, UpdateCellIp, Src/celllist.F90:307-410): a list model of head
  insertion and unlinking after random move sequences.
"""
import numpy as np
import pytest

import oracle


def _brute_pairs(ro, box, rl):
    ro = np.asarray(ro, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    d = ro[:, None, :] - ro[None, :, :]
    d -= box * np.rint(d / box)
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    i, j = np.nonzero(np.triu(r2 < rl * rl, k=1))
    return set(zip(i.tolist(), j.tolist()))


def _fluid(n, box, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.5, 0.5, (n, 3)) * np.asarray(box)


@pytest.mark.parametrize("box", [(40.0, 40.0, 40.0), (40.0, 33.0, 14.0), (11.0, 11.0, 11.0)])
def test_vlist_md_parity_half_list(box):
    rcut, skin = 5.0, 1.0
    ro = _fluid(700, box, 1)
    pairs = _brute_pairs(ro, box, rcut + skin)
    nn, rows = oracle.vlist("md", ro, box, rcut, skin)
    seen = set()
    for i, row in enumerate(rows):
        assert np.all(np.diff(row) > 0), "rows ascending"
        for j in row.tolist():
            p = (min(i, j), max(i, j))
            assert p not in seen, "each pair once"
            seen.add(p)
            # 1-based ip + jp odd: the lower index keeps the pair; even: the higher one
            lo, hi = p
            owner = lo if (lo + hi) % 2 == 1 else hi
            assert i == owner, (i, j)
    assert seen == pairs
    # the parity rule shares the pairs out about evenly
    assert abs(nn.sum() - len(pairs)) == 0


@pytest.mark.parametrize("box", [(40.0, 40.0, 40.0), (40.0, 33.0, 14.0)])
def test_vlist_mc_full_list(box):
    rcut, skin = 5.0, 1.0
    ro = _fluid(700, box, 2)
    pairs = _brute_pairs(ro, box, rcut + skin)
    _, rows = oracle.vlist("mc", ro, box, rcut, skin)
    want = [[] for _ in range(len(ro))]
    for i, j in sorted(pairs):
        want[i].append(j)
        want[j].append(i)
    for i in range(len(ro)):
        assert rows[i].tolist() == sorted(want[i])


def test_vlist_boundary_exact_distance():
    """r2 == rs2cut is not in the list (r2 < rs2cut), across the periodic boundary too."""
    box = (30.0, 30.0, 30.0)
    ro = np.array([[0.0, 0.0, 0.0], [6.0, 0.0, 0.0], [0.0, 5.999999, 0.0],
                   [-13.0, 10.0, 10.0], [14.0, 10.0, 10.0], [-12.0, -10.0, 0.0], [12.0, -10.0, 0.0]])
    for kind in ("md", "mc"):
        _, rows = oracle.vlist(kind, ro, box, 5.0, 1.0)
        listed = {(min(i, j), max(i, j)) for i, r in enumerate(rows) for j in r.tolist()}
        assert (0, 1) not in listed   # exactly rcut + skin apart
        assert (0, 2) in listed       # just inside
        assert (3, 4) in listed       # 3 A apart through the boundary
        assert (5, 6) not in listed   # exactly 6 A through the boundary


@pytest.mark.parametrize("kind", ["md", "mc"])
@pytest.mark.parametrize("box", [(60.0, 60.0, 60.0), (60.0, 45.0, 16.0)])
def test_vlist_cells_equal_brute(kind, box):
    ro = _fluid(3000, box, 3)
    nn, rows = oracle.vlist(kind, ro, box, 6.0, 1.5)
    nn2, rows2 = oracle.vlist(kind, ro, box, 6.0, 1.5, cells=True)
    assert np.array_equal(nn, nn2)
    assert all(np.array_equal(a, b) for a, b in zip(rows, rows2))


def test_vlist_overflow_raises():
    ro = _fluid(200, (10.0, 10.0, 10.0), 4)
    with pytest.raises(RuntimeError):
        oracle.vlist("mc", ro, (10.0, 10.0, 10.0), 4.0, 1.0, maxn=3)


@pytest.mark.parametrize("seq,m1,m2", [
    ([1.0, 2.0, 3.0], 9.0, 0.0),   # each value a new maximum: r2max2 never set
    ([3.0, 2.0, 1.0], 9.0, 4.0),
    ([2.0, 3.0, 1.0], 9.0, 1.0),   # the true second largest (4) is lost
    ([2.0, 2.0], 4.0, 4.0),        # a tie goes to r2max2
    ([0.0, 0.0], 0.0, 0.0),
])
def test_nlist_skin_update_order(seq, m1, m2):
    d = np.zeros((len(seq), 3))
    d[:, 0] = seq
    _, r1, r2 = oracle.nlist_skin(d, 100.0)
    assert (r1, r2) == (m1, m2)


def test_nlist_skin_decision():
    d = np.zeros((3, 3))
    d[:, 1] = [1.0, 2.0, 3.0]
    assert oracle.nlist_skin(d, 2.5)[0]           # 3 + 0 > 2.5
    assert not oracle.nlist_skin(d, 3.0)[0]       # 3 + 0 = 3: not above (two largest would give 5)
    d[:, 1] = [3.0, 2.0, 1.0]
    assert oracle.nlist_skin(d, 4.5)[0]           # 3 + 2 > 4.5
    assert not oracle.nlist_skin(d, 5.0)[0]
    # a displacement along all three axes: r2 = x^2 + y^2 + z^2
    d = np.array([[1.0, 2.0, 2.0]])
    assert oracle.nlist_skin(d, 100.0)[1] == 9.0


def test_linked_cell_list_updates():
    box = (50.0, 50.0, 50.0)
    rng = np.random.default_rng(5)
    n = 400
    ro = _fluid(n, box, 6)
    cl = oracle.LinkedCellList(ro, box, 10.0, 10.0)
    model = [list(m) for m in cl.members()]   # descending particle index per cell
    for m in model:
        assert m == sorted(m, reverse=True)
    cell = {i: c for c, m in enumerate(model) for i in m}
    moved = 0
    for _ in range(3000):
        ip = int(rng.integers(n))
        rnew = cl.ro[3 * ip:3 * ip + 3] + rng.uniform(-6.0, 6.0, 3)
        rnew = (rnew + 25.0) % 50.0 - 25.0
        changed = cl.update(ip, rnew)
        ci = np.floor((rnew + 25.0) * (1.0 / cl.cs)).astype(int)
        cnew = int(ci[0] + cl.nc[0] * (ci[1] + cl.nc[1] * ci[2]))
        assert changed == (cnew != cell[ip])
        if changed:
            moved += 1
            model[cell[ip]].remove(ip)
            model[cnew].insert(0, ip)
            cell[ip] = cnew
    assert moved > 1000
    assert [list(m) for m in cl.members()] == model
    assert np.array_equal(cl.npart, [len(m) for m in model])
    for c, m in enumerate(model):
        for k, i in enumerate(m):
            assert cl.prev[i] == (m[k - 1] if k else -1)
            assert cl.cellof[i] == c
    # membership equals a fresh SetCellList of the final positions
    fresh = oracle.LinkedCellList(cl.ro, box, 10.0, 10.0)
    assert [sorted(m) for m in fresh.members()] == [sorted(m) for m in model]
