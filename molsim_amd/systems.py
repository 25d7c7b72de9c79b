"""Seeded synthetic systems for the five workloads the path is measured and tested on.

The generators follow the semantics of MOLSIM's start-configuration code (uniform insertion
with overlap rejection, ``Src/coordinate.F90:1302-1344,2633-2716``; random-walk chains,
``Src/coordinate.F90:1609-1677``) but are this package's own code.  Units are MOLSIM's:
lengths in Angstrom, energies in kJ/mol (``Src/molsim.F90:770-835``).

Positions are returned in the box ``[-L/2, L/2)`` (MOLSIM's convention), shape (N, 3),
float64, row-major -- the same memory layout as Fortran ``ro(3,np)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# MOLSIM's constants, Src/mol.F90:70-76 and Src/molsim.F90:805.
ECH = 1.602176487e-19
EPSI0 = 8.854187817e-12
AVNO = 6.02214179e23
GAS_CONSTANT = 8.314472
EPSI0FOURPI = ECH ** 2 / (4.0 * math.pi * EPSI0 * 1.0e-10) * AVNO / 1.0e3  # kJ/mol * A

TEMPERATURE = 298.0


def beta_kjmol(temp: float = TEMPERATURE) -> float:
    """beta = sclene/(R T) in mol/kJ (Src/molsim.F90:252)."""
    return 1.0e3 / (GAS_CONSTANT * temp)


def relpermitt_from_bjerrum(lb: float, temp: float = TEMPERATURE) -> float:
    """Relative permittivity that gives Bjerrum length ``lb`` (A) at ``temp``."""
    return EPSI0FOURPI * beta_kjmol(temp) / lb


@dataclass
class SystemSpec:
    """Everything one energy/force evaluation needs, in MOLSIM's terms."""

    name: str
    ro: np.ndarray              # (N, 3) float64 in [-L/2, L/2)
    ipt: np.ndarray             # (N,) int32, 0-based particle type
    az: np.ndarray              # (N,) float64 charges
    box: np.ndarray             # (3,) float64
    pot: dict                   # potential description (see molsim_amd.engine.Potential)
    rcut: float
    drnlist: float = 0.0
    list_mode: str = "cell"     # "all" | "cell" | "verlet"
    ewald: str = "spm"          # "none" | "std" | "spm"
    alpha: float = 0.0
    ncut: int = 0
    ncutregion: str = "sphere"  # std Ewald k-vectors: "sphere" n^2 <= ncut^2, "cube" n^2 <= 3 ncut^2
    nmesh: int = 0
    order: int = 0
    lsurf: bool = True
    extra: dict = field(default_factory=dict)
    laz: np.ndarray | None = None  # weak charges: charge state per particle (az = z[type] or 0)

    @property
    def n(self) -> int:
        return int(self.az.size)

    @property
    def npt(self) -> int:
        return int(self.pot["npt"])


def _wrap(x, L):
    return (x + 0.5 * L) % L - 0.5 * L


def _insert_uniform(rng, n, L, dmin, existing=None, max_rounds=200):
    """Uniform insertion with overlap rejection (rejected particles are redrawn)."""
    from scipy.spatial import cKDTree

    pts = rng.uniform(0.0, L, size=(n, 3))
    fixed = None if existing is None or len(existing) == 0 else (existing + 0.5 * L) % L
    ftree = None if fixed is None or dmin <= 0 else cKDTree(fixed, boxsize=L)
    for _ in range(max_rounds):
        bad = np.zeros(n, dtype=bool)
        if dmin > 0:
            tree = cKDTree(pts, boxsize=L)
            pairs = tree.query_pairs(dmin, output_type="ndarray")
            if len(pairs):
                bad[pairs[:, 1]] = True
            if ftree is not None:
                hits = ftree.query_ball_point(pts, dmin, return_length=True)
                bad |= hits > 0
        if not bad.any():
            return pts - 0.5 * L
        pts[bad] = rng.uniform(0.0, L, size=(int(bad.sum()), 3))
    raise RuntimeError("overlap rejection did not converge")


def _random_walk_chains(rng, nchain, nbead, bond, L, dmin):
    """Self-avoiding random-walk chains built bead by bead with rejection."""
    cell = max(dmin, 1e-6)
    ncell = max(1, int(L // cell))
    cs = L / ncell
    grid: dict = {}
    out = np.zeros((nchain * nbead, 3))

    def cell_of(p):
        return tuple(((np.floor((p + 0.5 * L) / cs)).astype(int)) % ncell)

    def free(p):
        c = cell_of(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                    for q in grid.get(key, ()):
                        d = _wrap(p - q, L)
                        if d @ d < dmin * dmin:
                            return False
        return True

    k = 0
    for _ in range(nchain):
        while True:
            p = rng.uniform(-0.5 * L, 0.5 * L, size=3)
            if free(p):
                break
        start = k
        for b in range(nbead):
            if b > 0:
                for _try in range(1000):
                    v = rng.normal(size=3)
                    v *= bond / np.linalg.norm(v)
                    q = _wrap(out[k - 1] + v, L)
                    if free(q):
                        p = q
                        break
                else:
                    raise RuntimeError("chain growth failed")
            out[k] = p
            grid.setdefault(cell_of(p), []).append(p.copy())
            k += 1
        assert k - start == nbead
    return out


def _random_walk_chains_batch(rng, nchain, nbead, bond, L, dmin, max_tries=1000):
    """Self-avoiding random-walk chains grown all at once, one bead per chain per step;
    a proposal overlapping any placed bead or another proposal is redrawn."""
    from scipy.spatial import cKDTree

    out = np.zeros((nchain, nbead, 3))
    placed = np.zeros((0, 3))
    for b in range(nbead):
        todo = np.arange(nchain)
        prop = np.zeros((nchain, 3))
        ptree = cKDTree(placed + 0.5 * L, boxsize=L) if placed.shape[0] else None
        for _ in range(max_tries):
            if b == 0:
                q = rng.uniform(-0.5 * L, 0.5 * L, size=(todo.size, 3))
            else:
                v = rng.normal(size=(todo.size, 3))
                v *= bond / np.linalg.norm(v, axis=1)[:, None]
                q = _wrap(out[todo, b - 1] + v, L)
            prop[todo] = q
            bad = np.zeros(nchain, dtype=bool)
            if ptree is not None:
                hits = ptree.query_ball_point((q + 0.5 * L) % L, dmin, return_length=True)
                bad[todo[hits > 0]] = True
            pairs = cKDTree((prop + 0.5 * L) % L, boxsize=L).query_pairs(dmin, output_type="ndarray")
            if len(pairs):
                bad[np.maximum(pairs[:, 0], pairs[:, 1])] = True
            todo = np.nonzero(bad)[0]
            if todo.size == 0:
                break
        else:
            raise RuntimeError("chain growth failed")
        out[:, b] = prop
        placed = np.concatenate([placed, prop])
    return out.reshape(-1, 3)


def _pot(npt, z, form, rcut, relp, alpha, sig=None, eps=None, rad=None, lewald=True):
    return dict(
        npt=npt, form=form, lewald=lewald, alpha=alpha, rcut=rcut,
        epsi4pi=EPSI0FOURPI / relp,
        z=list(z), sig=list(sig or [0.0] * npt), eps=list(eps or [0.0] * npt),
        rad=list(rad or [0.0] * npt),
        utoltab=1e-5, ftoltab=1e-5, umaxtab=2e4, fmaxtab=2e4, r2uminin=0.1,
    )


def electrolyte_primitive(seed: int = 1, n: int = 1000, L: float = 80.0) -> SystemSpec:
    """1:1 primitive-model electrolyte: n/2 cations, n/2 anions, hard spheres d = 4 A,
    Bjerrum length 7.1 A, standard Ewald rc = 40 A, alpha = 0.08/A, |k| <= 10 (2 pi/L)."""
    rng = np.random.default_rng(seed)
    ro = _insert_uniform(rng, n, L, 4.0)
    ipt = np.array([0] * (n // 2) + [1] * (n - n // 2), dtype=np.int32)
    az = np.where(ipt == 0, 1.0, -1.0)
    relp = relpermitt_from_bjerrum(7.1)
    pot = _pot(2, [1.0, -1.0], 0, 40.0, relp, 0.08, rad=[2.0, 2.0])
    return SystemSpec("electrolyte_1k_stdewald", ro, ipt, az, np.array([L, L, L]), pot, 40.0,
                      list_mode="all", ewald="std", alpha=0.08, ncut=10, lsurf=True,
                      extra=dict(dtran=2.0, ntrial=10000))


def polyelectrolyte(seed: int = 2, nchain: int = 64, nbead: int = 256, L: float = 400.0,
                    nmesh: int = 128, order: int = 6, alpha: float = 0.15, rcut: float = 20.0,
                    lb: float = 7.1) -> SystemSpec:
    """64 chains x 256 beads (q = -1, d = 4 A, random walk with 5 A bonds) plus as many
    monovalent counterions, cubic box 400 A, SPME 128^3 order 6, alpha 0.15/A, rc 20 A
    cell list, LJ (sigma 4 A, eps 0.1 kT) + Coulomb table."""
    rng = np.random.default_rng(seed)
    beads = _random_walk_chains(rng, nchain, nbead, 5.0, L, 4.0)
    ions = _insert_uniform(rng, nchain * nbead, L, 4.0, existing=beads)
    ro = np.concatenate([beads, ions])
    nb = nchain * nbead
    ipt = np.array([0] * nb + [1] * nb, dtype=np.int32)
    az = np.where(ipt == 0, -1.0, 1.0)
    relp = relpermitt_from_bjerrum(lb)
    eps = 0.1 / beta_kjmol()
    pot = _pot(2, [-1.0, 1.0], 1, rcut, relp, alpha, sig=[4.0, 4.0], eps=[eps, eps],
               rad=[2.0, 2.0])
    return SystemSpec(f"polyelectrolyte_{ro.shape[0]}_spme{nmesh}", ro, ipt, az,
                      np.array([L, L, L]), pot, rcut, list_mode="cell", ewald="spm",
                      alpha=alpha, nmesh=nmesh, order=order, lsurf=True,
                      extra=dict(nchain=nchain, nbead=nbead, bond_r0=5.0,
                                 bond_k=0.4 / beta_kjmol()))


def ionic_fluid(seed: int = 3, n: int = 1 << 20, rho: float = 0.002, nmesh: int = 256,
                order: int = 6, rcut: float = 18.0, skin: float = 2.0, alpha: float = 0.18,
                lb: float = 7.1) -> SystemSpec:
    """Coarse-grained +-1 ionic fluid, LJ sigma 4 A eps 0.1 kT, number density 0.002/A^3,
    SPME 256^3 order 6, rc 18 A with a 2 A Verlet skin."""
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1.0 / 3.0)
    ro = _insert_uniform(rng, n, L, 3.6)
    ipt = (np.arange(n) % 2).astype(np.int32)
    az = np.where(ipt == 0, 1.0, -1.0)
    relp = relpermitt_from_bjerrum(lb)
    eps = 0.1 / beta_kjmol()
    pot = _pot(2, [1.0, -1.0], 1, rcut, relp, alpha, sig=[4.0, 4.0], eps=[eps, eps])
    return SystemSpec(f"ionic_fluid_{n}_spme{nmesh}", ro, ipt, az, np.array([L, L, L]), pot, rcut,
                      drnlist=skin, list_mode="verlet", ewald="spm", alpha=alpha, nmesh=nmesh,
                      order=order, lsurf=False)


def small_spme_system(seed: int = 5, n: int = 400, L: float = 60.0, nmesh: int = 32,
                      order: int = 6, alpha: float = 0.3, rcut: float = 12.0, form: int = 1,
                      charges=(1.0, -1.0), lb: float = 7.1, lsurf=True, list_mode="cell",
                      ) -> SystemSpec:
    """Small electrolyte for parity tests at sizes the oracle finishes in a second."""
    rng = np.random.default_rng(seed)
    ro = _insert_uniform(rng, n, L, 3.0)
    nt = len(charges)
    ipt = (np.arange(n) % nt).astype(np.int32)
    az = np.asarray(charges, dtype=np.float64)[ipt]
    relp = relpermitt_from_bjerrum(lb)
    eps = 0.1 / beta_kjmol()
    pot = _pot(nt, charges, form, rcut, relp, alpha, sig=[3.0] * nt, eps=[eps] * nt,
               rad=[1.5] * nt)
    return SystemSpec(f"small_{n}", ro, ipt, az, np.array([L, L, L]), pot, rcut,
                      list_mode=list_mode, ewald="spm", alpha=alpha, nmesh=nmesh, order=order,
                      lsurf=lsurf)


# Four ion types shared by every member of an ensemble, so that one table serves all of them:
# 0: +1, 1: -1 (also the chain beads), 2: +2, 3: -2.
ENSEMBLE_CHARGES = (1.0, -1.0, 2.0, -2.0)


def ensemble_mesh(L: float, hmax: float = 6.5, kmin: int = 16, kmax: int = 64) -> int:
    """SPME mesh for box length L: the smallest power of two with spacing L/K <= hmax,
    clamped to [kmin, kmax] (16^3 at L = 87.5 A up to 64^3 at L = 406 A)."""
    k = kmin
    while k < kmax and L / k > hmax:
        k *= 2
    return k


def ensemble_member(seed: int, n: int, kind: str, f2: float, phi: float = 0.01,
                    rcut: float = 24.0, alpha: float = 0.125, order: int = 6, lb: float = 7.1,
                    nbead: int = 32) -> SystemSpec:
    """One ensemble member: n particles of diameter 4 A at volume fraction phi, SPME order 6,
    rc 24 A with alpha 0.125/A (alpha rc = 3), LJ (sigma 4 A, eps 0.1 kT) + hard core +
    Coulomb at Bjerrum length 7.1 A.  ``kind`` "electrolyte": +-1 and +-2 ion pairs, a share
    f2 of the pairs divalent; "polyelectrolyte": chains of nbead -1 beads (5 A random-walk
    bonds) neutralised by +1 and +2 counterions, a share f2 of the counter-charge divalent."""
    rng = np.random.default_rng(seed)
    L = (n * (math.pi / 6.0) * 4.0 ** 3 / phi) ** (1.0 / 3.0)
    if kind == "electrolyte":
        npair = n // 2
        n2 = int(round(f2 * npair))
        n1 = npair - n2
        ipt = np.array([0] * n1 + [1] * n1 + [2] * n2 + [3] * n2, dtype=np.int32)
        ro = _insert_uniform(rng, ipt.size, L, 4.0)
    else:
        nb = max(nbead, int(math.ceil(n / (2.0 - 0.5 * f2) / nbead)) * nbead)
        nch = nb // nbead
        n2 = int(round(0.5 * f2 * nb))
        n1 = nb - 2 * n2
        beads = _random_walk_chains_batch(rng, nch, nbead, 5.0, L, 4.0)
        ions = _insert_uniform(rng, n1 + n2, L, 4.0, existing=beads)
        ro = np.concatenate([beads, ions])
        ipt = np.array([1] * nb + [0] * n1 + [2] * n2, dtype=np.int32)
    az = np.asarray(ENSEMBLE_CHARGES)[ipt]
    assert abs(az.sum()) < 1e-9
    relp = relpermitt_from_bjerrum(lb)
    eps = 0.1 / beta_kjmol()
    pot = _pot(4, ENSEMBLE_CHARGES, 1, rcut, relp, alpha, sig=[4.0] * 4, eps=[eps] * 4,
               rad=[2.0] * 4)
    K = ensemble_mesh(L)
    return SystemSpec(f"{kind}_{ipt.size}_spme{K}", ro, ipt, az, np.array([L, L, L]), pot, rcut,
                      list_mode="cell", ewald="spm", alpha=alpha, nmesh=K, order=order,
                      lsurf=True, extra=dict(kind=kind, f2=f2, dtran=2.0, ntrial=1000))


def _member_args(a):
    seed, n, kind, f2, kw = a
    return ensemble_member(seed, n, kind, f2, **kw)


def ensemble(nsys: int = 512, seed: int = 4, nmin: int = 200, nmax: int = 20000,
             workers: int = 1, **kw) -> list[SystemSpec]:
    """Independent electrolyte / polyelectrolyte systems with particle counts log-uniform in
    [nmin, nmax], box at volume fraction 0.01, mesh 16^3..64^3 by box length, seeded
    shares of divalent ions.  ``workers`` > 1 generates in forked processes (call it before
    the process touches the GPU)."""
    rng = np.random.default_rng(seed)
    logn = rng.uniform(math.log(nmin), math.log(nmax), size=nsys)
    kinds = rng.choice(["electrolyte", "polyelectrolyte"], size=nsys)
    f2 = rng.uniform(0.0, 0.5, size=nsys)
    seeds = rng.integers(1, 2 ** 31 - 1, size=nsys)
    args = [(int(seeds[i]), max(nmin, int(round(math.exp(logn[i]) / 2)) * 2), str(kinds[i]),
             float(f2[i]), kw) for i in range(nsys)]
    if workers > 1:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_member_args, args, chunksize=4)
    return [_member_args(a) for a in args]


def served_system(seed: int, n: int, nbead_max: int = 256, order: int = 6, alpha: float = 0.15,
                  rcut: float = 20.0, lb: float = 7.1, h: float = 3.125) -> SystemSpec:
    """The polyelectrolyte composition rescaled to n particles: half of them -1 beads in
    random-walk chains of up to 256 beads (5 A bonds), half +1 counterions, at the 400 A /
    32768-particle density (box 400 (n / 32768)^(1/3) A); SPME order 6 with the mesh spacing
    of 128^3 in 400 A (K the smallest power of two with L/K <= 3.125 A, 16..256), alpha
    0.15/A, rc 20 A cell list, LJ (sigma 4 A, eps 0.1 kT) + hard core + Coulomb table."""
    rng = np.random.default_rng(seed)
    L = 400.0 * (n / 32768.0) ** (1.0 / 3.0)
    nb_total = n // 2
    nchain = max(1, int(math.ceil(nb_total / nbead_max)))
    nbead = max(1, nb_total // nchain)
    nb = nchain * nbead
    beads = _random_walk_chains_batch(rng, nchain, nbead, 5.0, L, 4.0)
    ions = _insert_uniform(rng, nb, L, 4.0, existing=beads)
    ro = np.concatenate([beads, ions])
    ipt = np.array([0] * nb + [1] * nb, dtype=np.int32)
    az = np.where(ipt == 0, -1.0, 1.0)
    relp = relpermitt_from_bjerrum(lb)
    eps = 0.1 / beta_kjmol()
    pot = _pot(2, [-1.0, 1.0], 1, rcut, relp, alpha, sig=[4.0, 4.0], eps=[eps, eps], rad=[2.0, 2.0])
    K = 16
    while K < 256 and L / K > h:
        K *= 2
    return SystemSpec(f"served_{ro.shape[0]}_spme{K}", ro, ipt, az, np.array([L, L, L]), pot, rcut,
                      list_mode="cell", ewald="spm", alpha=alpha, nmesh=K, order=order, lsurf=True)


def served_pool(nsys: int = 48, seed: int = 5, nmin: int = 100, nmax: int = 100000,
                workers: int = 1) -> dict:
    """{sys_id: spec} with particle counts log-uniform in [nmin, nmax]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nsys))
    seeds = rng.integers(1, 2 ** 31 - 1, size=nsys)
    args = [(int(seeds[i]), max(nmin, int(round(ns[i] / 2)) * 2)) for i in range(nsys)]
    if workers > 1:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(workers) as pool:
            specs = pool.starmap(served_system, args, chunksize=2)
    else:
        specs = [served_system(*a) for a in args]
    return dict(enumerate(specs))
