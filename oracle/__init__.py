"""ORACLE -- test infrastructure only.

FP64 CPU restatement of MOLSIM's non-bonded energy/force path (``oracle.cpp``) behind a
numpy-friendly ctypes layer.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the checker; the
product (``molsim_amd``) never imports it.

Pinning status of each function (how it is checked; see tests/test_oracle_*.py):

* bspline        -- bit for bit against the reference's compiled CardinalBSpline
                    (oracle/_ref, tests/test_oracle_ref_mollib.py); partition of unity,
                    symmetry, closed-form cubic, finite-difference derivatives.
* build_table    -- tabulated U and F against the analytic potential within utoltab/ftoltab
                    (the reference's TestUBuffer semantics) plus C0/C1 continuity at joints.
* cell lists     -- brute-force O(N^2) enumeration (the reference's TestCellList check).
* fft3           -- numpy.fft.rfftn / irfftn on the same arrays.
* spme_setup     -- b^2 moduli from numpy's FFT of the spline, closed-form influence function.
* std Ewald      -- NaCl Madelung constant, and agreement of std Ewald with SPME at high
                    order / fine mesh (the reference's TestEwaldSPM recipe).
* utotal forces  -- finite differences of the oracle's own energy, translation invariance.
* dutotal        -- identity dU = U(new) - U(old) from two full evaluations (the reference's
                    drift check).
* random         -- bit for bit against the reference's compiled Random (oracle/_ref: 10^5
                    draws and the end state for several seeds; committed vectors in
                    tests/golden/ref_mollib_vectors.npz).
* std Ewald cube -- k-vector counts of both regions by enumeration; cube and sphere agree
                    once both hold the converged sum.
* vlist          -- the set of pairs within rcut + drnlist by numpy brute force (a different
                    minimum-image formula); each pair exactly once in the MD half lists, in the
                    row the parity rule names; MC rows = symmetric closure; a pair exactly at
                    rcut + drnlist is excluded (r2 < rs2cut); cell-list form = O(N^2) form.
* nlist_skin     -- hand-computed r2max1/r2max2 for displacement sequences where the
                    reference's update order differs from the true two largest values.
* celllist update-- a Python list model of head insertion / unlinking after random move
                    sequences; prev/next consistency; membership equals a fresh SetCellList.

oracle/_ref/libref_mollib.so (``make -C oracle ref``) is the reference's Src/mollib.F90
compiled where it lies with amdflang, behind the wrapper ref_mollib.F90.  No other file of
the reference's path builds here (Src/mol.F90 fails to compile with amdflang).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile oracle.cpp with the committed Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_LIB_OMP_PATH = os.path.join(_HERE, "liboracle_omp.so")
_lib_omp = None


def lib_omp():
    """The same oracle built with OpenMP (pair loop, FFT lines, gather and list rows shared out
    over the host's threads): the CPU baseline on all cores, and the checker only where the
    single-threaded build would take minutes (the million-particle system); a CPU test checks
    that both builds agree."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(_LIB_OMP_PATH):
            subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle_omp.so"], check=True)
        _lib_omp = C.CDLL(_LIB_OMP_PATH)
        _declare(_lib_omp)
    return _lib_omp


def threads(omp: bool = False) -> int:
    """Threads the oracle's loops use (1 for the checking oracle)."""
    return int((lib_omp() if omp else lib()).orc_threads())


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcPot(C.Structure):
    _fields_ = [
        ("npt", C.c_int), ("form", C.c_int), ("lewald", C.c_int), ("pad", C.c_int),
        ("alpha", C.c_double), ("rcut", C.c_double), ("epsi4pi", C.c_double),
        ("utoltab", C.c_double), ("ftoltab", C.c_double), ("umaxtab", C.c_double),
        ("fmaxtab", C.c_double), ("r2uminin", C.c_double),
        ("z", C.c_double * 8), ("sig", C.c_double * 8), ("eps", C.c_double * 8),
        ("rad", C.c_double * 8),
    ]


class OrcSys(C.Structure):
    _fields_ = [
        ("np", C.c_int), ("npt", C.c_int),
        ("ro", C.POINTER(C.c_double)), ("ipt", C.POINTER(C.c_int)), ("az", C.POINTER(C.c_double)),
        ("box", C.c_double * 3),
        ("rcut", C.c_double), ("drnlist", C.c_double), ("epsi4pi", C.c_double),
        ("ubuf", C.POINTER(C.c_double)), ("nbuf", C.c_int),
        ("iubuflow", C.POINTER(C.c_int)), ("r2umin", C.POINTER(C.c_double)),
        ("r2atat", C.POINTER(C.c_double)),
        ("list", C.c_int), ("ewald", C.c_int), ("surf", C.c_int),
        ("alpha", C.c_double), ("ncut", C.c_int), ("region", C.c_int),
        ("K", C.c_int), ("order", C.c_int),
        ("laz", C.POINTER(C.c_int)), ("zat", C.POINTER(C.c_double)),
    ]


def _declare(L):
    L.orc_bspline.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_int), _dp, _dp]
    L.orc_random.argtypes = [C.POINTER(C.c_int32), _ip, C.POINTER(C.c_double)]
    L.orc_random.restype = C.c_double
    L.orc_heapsort_index.argtypes = [C.c_int, _dp, _ip]
    L.orc_pot_eval.argtypes = [C.POINTER(OrcPot), C.c_int, C.c_int, C.c_double,
                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_build_table.argtypes = [C.POINTER(OrcPot), _dp, C.c_int, _ip, _ip, _dp, _dp]
    L.orc_build_table.restype = C.c_int
    L.orc_cell_init.argtypes = [_dp, C.c_double, C.c_double, _ip, _dp, _ip, C.c_int]
    L.orc_cell_init.restype = C.c_int
    L.orc_cell_build.argtypes = [C.c_int, _dp, _dp, _ip, _dp, _ip, _ip, _ip, _ip]
    L.orc_cell_build.restype = C.c_int
    L.orc_ewald_std_kfac.argtypes = [_dp, C.c_double, C.c_int, C.c_int, _dp, C.c_int]
    L.orc_ewald_std_kfac.restype = C.c_int
    L.orc_ewald_std_rec.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                    C.c_double, _dp, C.POINTER(C.c_double), _dp,
                                    C.POINTER(C.c_double), C.c_void_p]
    L.orc_ewald_std_rec.restype = C.c_int
    L.orc_fft3_r2c.argtypes = [C.c_int, _dp, _dp]
    L.orc_fft3_c2r.argtypes = [C.c_int, _dp, _dp]
    L.orc_spme_setup.argtypes = [_dp, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp]
    L.orc_spme_setup.restype = C.c_int
    L.orc_spme_spread.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, _dp]
    L.orc_spme_conv.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    L.orc_spme_energy.argtypes = [C.c_int, _dp, _dp]
    L.orc_spme_energy.restype = C.c_double
    L.orc_spme_gather.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp, _dp]
    L.orc_ewald_self.argtypes = [C.c_int, _dp, C.c_double]
    L.orc_ewald_self.restype = C.c_double
    L.orc_ewald_surf.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_void_p, C.c_void_p]
    L.orc_ewald_surf.restype = C.c_double
    L.orc_utotal.argtypes = [C.POINTER(OrcSys), _dp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                             _dp, C.POINTER(C.c_double)]
    L.orc_utotal.restype = C.c_int
    L.orc_dutotal.argtypes = [C.POINTER(OrcSys), C.c_int, _dp, C.POINTER(C.c_int), _dp,
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_dutotal.restype = C.c_int
    L.orc_dutotal_charge.argtypes = [C.POINTER(OrcSys), C.c_int, _ip, C.POINTER(C.c_int), _dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_dutotal_charge.restype = C.c_int
    for f in (L.orc_vlist, L.orc_vlist_cells):
        f.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, _ip, _ip]
        f.restype = C.c_int
    L.orc_nlist_skin.argtypes = [C.c_int, _dp, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_nlist_skin.restype = C.c_int
    L.orc_celllist_set.argtypes = [C.c_int, _dp, _dp, _ip, _dp, _ip, _ip, _ip, _ip, _ip]
    L.orc_celllist_set.restype = C.c_int
    L.orc_celllist_update.argtypes = [C.c_int, _dp, _dp, _ip, _dp, _ip, _ip, _ip, _ip, _ip]
    L.orc_celllist_update.restype = C.c_int
    L.orc_trials_open.argtypes = [C.POINTER(OrcSys)]
    L.orc_trials_open.restype = C.c_void_p
    L.orc_trials_close.argtypes = [C.c_void_p]
    L.orc_trials_close.restype = None
    L.orc_trials_dutotal.argtypes = [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_int), _dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_trials_dutotal.restype = C.c_int


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")
    return rc


# ---------------------------------------------------------------------------------------
def bspline(order: int, pos: float):
    vals = np.zeros(order)
    ders = np.zeros(order)
    k = C.c_int()
    lib().orc_bspline(order, float(pos), C.byref(k), vals, ders)
    return k.value, vals, ders


class Rng:
    """The reference's Random(iseed) generator state."""

    def __init__(self, iseed: int):
        self.idum = C.c_int32(iseed)
        self.state = np.array([-1, -1], dtype=np.int32)
        self.am = C.c_double(0.0)

    def __call__(self) -> float:
        return lib().orc_random(C.byref(self.idum), self.state, C.byref(self.am))


def heapsort_index(vec):
    vec = np.ascontiguousarray(vec, dtype=np.float64)
    idx = np.zeros(len(vec), dtype=np.int32)
    lib().orc_heapsort_index(len(vec), vec, idx)
    return idx


def make_pot(pot: dict) -> OrcPot:
    p = OrcPot()
    p.npt = pot["npt"]
    p.form = pot.get("form", 0)
    p.lewald = int(pot.get("lewald", 1))
    p.alpha = pot.get("alpha", 0.0)
    p.rcut = pot["rcut"]
    p.epsi4pi = pot["epsi4pi"]
    p.utoltab = pot.get("utoltab", 1e-5)
    p.ftoltab = pot.get("ftoltab", 1e-5)
    p.umaxtab = pot.get("umaxtab", 2e4)
    p.fmaxtab = pot.get("fmaxtab", 2e4)
    p.r2uminin = pot.get("r2uminin", 0.1)
    for k in ("z", "sig", "eps", "rad"):
        arr = getattr(p, k)
        for i, v in enumerate(pot.get(k, [0.0] * p.npt)):
            arr[i] = v
    return p


def pot_eval(pot: dict, it: int, jt: int, r: float):
    p = make_pot(pot)
    u0, u1, u2 = C.c_double(), C.c_double(), C.c_double()
    lib().orc_pot_eval(C.byref(p), it, jt, float(r), C.byref(u0), C.byref(u1), C.byref(u2))
    return u0.value, u1.value, u2.value


def build_table(pot: dict, maxbuf: int = 200000):
    p = make_pot(pot)
    npt = pot["npt"]
    nptpt = npt * (npt + 1) // 2
    ubuf = np.zeros(maxbuf)
    low = np.zeros(nptpt, dtype=np.int32)
    ng = np.zeros(nptpt, dtype=np.int32)
    r2umin = np.zeros(nptpt)
    r2atat = np.zeros(nptpt)
    n = _check(lib().orc_build_table(C.byref(p), ubuf, maxbuf, low, ng, r2umin, r2atat), "table")
    return dict(ubuf=ubuf[:n].copy(), iubuflow=low, nugrid=ng, r2umin=r2umin, r2atat=r2atat)


def cell_init(box, rcell, rcut):
    box = np.asarray(box, dtype=np.float64)
    nc = np.zeros(3, dtype=np.int32)
    cs = np.zeros(3)
    st = np.zeros(3 * 4096, dtype=np.int32)
    n = _check(lib().orc_cell_init(box, rcell, rcut, nc, cs, st, 4096), "cell_init")
    return nc, cs, st[: 3 * n].reshape(n, 3).copy()


def cell_build(ro, box, nc, cs):
    ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1)
    n = ro.size // 3
    ncell = int(np.prod(nc))
    head = np.zeros(ncell, dtype=np.int32)
    nxt = np.zeros(n, dtype=np.int32)
    npart = np.zeros(ncell, dtype=np.int32)
    cellof = np.zeros(n, dtype=np.int32)
    _check(lib().orc_cell_build(n, ro, np.asarray(box, dtype=np.float64), np.asarray(nc, dtype=np.int32),
                                np.asarray(cs, dtype=np.float64), head, nxt, npart, cellof),
           "cell_build")
    return head, nxt, npart, cellof


def cell_members(head, nxt):
    """Per cell, the particle order a linked-list traversal visits (iphead, ipnext, ...)."""
    out = []
    for h in head:
        lst = []
        j = int(h)
        while j >= 0:
            lst.append(j)
            j = int(nxt[j])
        out.append(lst)
    return out


VLIST_KIND = {"md": 0, "mc": 1}


def vlist(kind, ro, box, rcut, drnlist, maxn=None, cells=False, table=False, omp=False):
    """Verlet lists as SetVListMD (kind "md": parity half list) or SetVListMC ("mc": full list)
    build them: (nneighpn (N,), rows) with rows[i] the 0-based partners of particle i in
    ascending order.  cells=True takes the cell-list form (same rows, O(N)); with omp=True its
    rows are shared out over the host's cores (each row is formed the same way).  table=True
    returns rows as one (N, maxn) array with -1 past each row's end (large systems)."""
    ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1)
    n = ro.size // 3
    if maxn is None:
        maxn = max(n - 1, 1) if not cells else 512
    nn = np.zeros(n, dtype=np.int32)
    jl = np.zeros(n * maxn, dtype=np.int32)
    L = lib_omp() if omp else lib()
    f = L.orc_vlist_cells if cells else L.orc_vlist
    _check(f(VLIST_KIND[kind], n, ro, np.asarray(box, dtype=np.float64), float(rcut), float(drnlist), maxn, nn, jl),
           "vlist")
    jl = jl.reshape(n, maxn)
    if table:
        jl[np.arange(maxn)[None, :] >= nn[:, None]] = -1
        return nn, jl
    return nn, [jl[i, :nn[i]].copy() for i in range(n)]


def nlist_skin(drosum, drnlist):
    """NList skin test (inlist = 0): (rebuild, r2max1, r2max2) from the displacements summed
    since the last build, drosum (N, 3)."""
    d = np.ascontiguousarray(drosum, dtype=np.float64).reshape(-1)
    m1, m2 = C.c_double(), C.c_double()
    rb = lib().orc_nlist_skin(d.size // 3, d, float(drnlist), C.byref(m1), C.byref(m2))
    return bool(rb), m1.value, m2.value


class LinkedCellList:
    """MOLSIM's cell list with its doubly linked per-cell lists (SetCellList, UpdateCellIp)."""

    def __init__(self, ro, box, rcell, rcut):
        self.box = np.asarray(box, dtype=np.float64)
        self.nc, self.cs, _ = cell_init(self.box, rcell, rcut)
        self.ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1).copy()
        n = self.ro.size // 3
        ncell = int(np.prod(self.nc))
        self.head = np.zeros(ncell, dtype=np.int32)
        self.npart = np.zeros(ncell, dtype=np.int32)
        self.next = np.zeros(n, dtype=np.int32)
        self.prev = np.zeros(n, dtype=np.int32)
        self.cellof = np.zeros(n, dtype=np.int32)
        _check(lib().orc_celllist_set(n, self.ro, self.box, self.nc, self.cs, self.head, self.npart,
                                      self.next, self.prev, self.cellof), "celllist_set")

    def update(self, ip, rnew):
        """MCUpdate's ro(ip) <- rnew followed by UpdateCellIp(ip); True when the cell changed."""
        self.ro[3 * ip:3 * ip + 3] = np.asarray(rnew, dtype=np.float64)
        return bool(_check(lib().orc_celllist_update(int(ip), self.ro, self.box, self.nc, self.cs, self.head,
                                                     self.npart, self.next, self.prev, self.cellof),
                           "celllist_update"))

    def members(self):
        """Per cell, the order a traversal from iphead along ipnext visits."""
        return cell_members(self.head, self.next)


def fft3_r2c(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    K = q.shape[0]
    out = np.zeros(2 * (K // 2 + 1) * K * K)
    lib().orc_fft3_r2c(K, q.reshape(-1), out)
    return out.view(np.complex128).reshape(K, K, K // 2 + 1)


def fft3_c2r(f, K):
    f = np.ascontiguousarray(f, dtype=np.complex128)
    out = np.zeros(K * K * K)
    lib().orc_fft3_c2r(K, f.view(np.float64).reshape(-1), out)
    return out.reshape(K, K, K)


def spme_setup(box, alpha, K, order):
    H = K // 2 + 1
    mf, ef, vf = np.zeros(H ** 3), np.zeros(H ** 3), np.zeros(H ** 3)
    _check(lib().orc_spme_setup(np.asarray(box, dtype=np.float64), alpha, K, order, mf, ef, vf),
           "spme_setup")
    return mf, ef, vf


def spme_spread(ro, az, box, K, order):
    ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1)
    az = np.ascontiguousarray(az, dtype=np.float64)
    Q = np.zeros(K ** 3)
    lib().orc_spme_spread(az.size, ro, az, np.asarray(box, dtype=np.float64), K, order, Q)
    return Q


def ewald_std_kfac(box, alpha, ncut, region=0):
    kf = np.zeros(4 * (ncut + 1) ** 3)
    n = _check(lib().orc_ewald_std_kfac(np.asarray(box, dtype=np.float64), alpha, ncut, region, kf,
                                        kf.size), "kfac")
    return kf[:n].copy()


class System:
    """Owns the arrays an OrcSys points at."""

    def __init__(self, ro, ipt, az, box, rcut, epsi4pi, table, *, npt, list_mode=1, drnlist=0.0,
                 ewald=2, surf=True, alpha=0.0, ncut=0, region=0, K=0, order=0, laz=None, zat=None):
        """laz (N,) charge states and zat (npt,) full charges per type switch on weak charges;
        az must then hold zat[type] where laz is set and 0 elsewhere, as MOLSIM keeps it."""
        self.ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1)
        self.ipt = np.ascontiguousarray(ipt, dtype=np.int32)
        self.az = np.ascontiguousarray(az, dtype=np.float64)
        self.table = {k: np.ascontiguousarray(v) for k, v in table.items()}
        s = OrcSys()
        s.np = self.az.size
        s.npt = npt
        s.ro = self.ro.ctypes.data_as(C.POINTER(C.c_double))
        s.ipt = self.ipt.ctypes.data_as(C.POINTER(C.c_int))
        s.az = self.az.ctypes.data_as(C.POINTER(C.c_double))
        for k in range(3):
            s.box[k] = float(box[k])
        s.rcut, s.drnlist, s.epsi4pi = rcut, drnlist, epsi4pi
        t = self.table
        s.ubuf = t["ubuf"].ctypes.data_as(C.POINTER(C.c_double))
        s.nbuf = t["ubuf"].size
        s.iubuflow = t["iubuflow"].ctypes.data_as(C.POINTER(C.c_int))
        s.r2umin = t["r2umin"].ctypes.data_as(C.POINTER(C.c_double))
        s.r2atat = t["r2atat"].ctypes.data_as(C.POINTER(C.c_double))
        s.list, s.ewald, s.surf = list_mode, ewald, int(surf)
        s.alpha, s.ncut, s.region, s.K, s.order = alpha, ncut, region, K, order
        self.laz = None if laz is None else np.ascontiguousarray(laz, dtype=np.int32)
        self.zat = None if zat is None else np.ascontiguousarray(zat, dtype=np.float64)
        if self.laz is not None:
            s.laz = self.laz.ctypes.data_as(C.POINTER(C.c_int))
            s.zat = self.zat.ctypes.data_as(C.POINTER(C.c_double))
        self.s = s
        self.npt = npt

    def set_positions(self, ro):
        self.ro[:] = np.asarray(ro, dtype=np.float64).reshape(-1)

    def utotal(self, omp: bool = False):
        """UTotal; omp=True runs the OpenMP build (multi-core timing only)."""
        nptpt = self.npt * (self.npt + 1) // 2
        ut = np.zeros(nptpt + 1)
        f = np.zeros(self.ro.size)
        urec, utot, vir = C.c_double(), C.c_double(), C.c_double()
        L = lib_omp() if omp else lib()
        _check(L.orc_utotal(C.byref(self.s), ut, C.byref(urec), C.byref(utot), f, C.byref(vir)),
               "utotal")
        return dict(u_twob=ut, u_rec=urec.value, u_tot=utot.value, force=f.reshape(-1, 3),
                    virial=vir.value)

    def dutotal(self, ip, rnew):
        nptpt = self.npt * (self.npt + 1) // 2
        dt = np.zeros(nptpt + 1)
        ov = C.c_int()
        drec, dtot = C.c_double(), C.c_double()
        _check(lib().orc_dutotal(C.byref(self.s), int(ip), np.asarray(rnew, dtype=np.float64), C.byref(ov),
                                 dt, C.byref(drec), C.byref(dtot)), "dutotal")
        return dict(overlap=bool(ov.value), du_twob=dt, du_rec=drec.value, du_tot=dtot.value)

    def trials(self):
        """Trial energies with the configuration's cell list, structure factors (or charge
        mesh) and dipole resident, as MOLSIM's MC loop keeps them: Trials.dutotal(ip, rnew)
        costs O(partners + nkvec) with std Ewald.  The positions must not change while it is
        open."""
        return Trials(self)

    def dutotal_charge(self, ips):
        """ChargeChange trial: the charge states of particles ips (one or two, 0-based) flip."""
        nptpt = self.npt * (self.npt + 1) // 2
        dt = np.zeros(nptpt + 1)
        ov = C.c_int()
        drec, dtot = C.c_double(), C.c_double()
        ia = np.ascontiguousarray(np.atleast_1d(ips), dtype=np.int32)
        _check(lib().orc_dutotal_charge(C.byref(self.s), int(ia.size), ia, C.byref(ov), dt,
                                        C.byref(drec), C.byref(dtot)), "dutotal_charge")
        return dict(overlap=bool(ov.value), du_twob=dt, du_rec=drec.value, du_tot=dtot.value)


class Trials:
    """Incremental DUTotal (orc_trials_*) over one System's configuration."""

    def __init__(self, system):
        self.sys = system
        self.h = lib().orc_trials_open(C.byref(system.s))
        if not self.h:
            raise RuntimeError("oracle trials_open failed")

    def dutotal(self, ip, rnew):
        nptpt = self.sys.npt * (self.sys.npt + 1) // 2
        dt = np.zeros(nptpt + 1)
        ov = C.c_int()
        drec, dtot = C.c_double(), C.c_double()
        _check(lib().orc_trials_dutotal(self.h, int(ip), np.asarray(rnew, dtype=np.float64), C.byref(ov),
                                        dt, C.byref(drec), C.byref(dtot)), "trials_dutotal")
        return dict(overlap=bool(ov.value), du_twob=dt, du_rec=drec.value, du_tot=dtot.value)

    def close(self):
        if self.h:
            lib().orc_trials_close(self.h)
            self.h = None

    def __del__(self):
        self.close()
