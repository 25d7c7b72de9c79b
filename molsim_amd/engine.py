"""Host-side mirror of MOLSIM's non-bonded call surface over the HIP C-ABI.

``Engine`` wraps one ``msg_ctx`` (``include/molsim_gpu.h``) and exposes the reference's
operations under the reference's names:

* ``utotal()``        -- UTotal(iStage)               Src/energy.F90:280-434
* ``dutotal()``       -- DUTotal(lhsoverlap, ...)     Src/denergy.F90:98-279
* ``mc_update()``     -- MCUpdate                      Src/mc.F90:5203-5235
* ``ewald_setup()``   -- EwaldSetup                    Src/energy.F90:6601-7126
* ``set_box()``       -- SetBoxParam + EwaldSetup + InitCellList (VolChange) Src/mc.F90:3423-3436
* ``nlist()``         -- NList / SetCellList           Src/nlist.F90:553-669
* ``build_table()``   -- PotTwoBodyTab1/2/3            Src/potential.F90:868-1258

Errors from the library raise ``MolsimGpuError`` carrying the C status code, in place of
the reference's ``Stop``.  There is no CPU fallback: without the built HIP library every
call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmolsim_gpu.so")
_lib = None

LIST_MODES = {"all": 0, "cell": 1, "verlet": 2}
EWALD_MODES = {"none": 0, "std": 1, "spm": 2}
PRECISIONS = {"mixed": 0, "fp64": 1}
MAX_TYPES = 8          # MSG_MAX_TYPES of the C-ABI
KERNELS = ["bin", "scan", "pair", "spread", "fft_x", "fft_y", "fft_zconv", "fft_yi",
           "fft_xi", "gather", "reduce", "ewald_std", "dutotal", "vlist", "scatter", "finalize"]


class MolsimGpuError(RuntimeError):
    def __init__(self, code, text):
        super().__init__(f"molsim_gpu error {code}: {text}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("np", C.c_int), ("npt", C.c_int), ("boxlen", C.c_double * 3),
        ("rcut", C.c_double), ("drnlist", C.c_double), ("epsi4pi", C.c_double),
        ("list_mode", C.c_int), ("ewald_mode", C.c_int), ("lsurf", C.c_int),
        ("alpha", C.c_double), ("ncut", C.c_int), ("ncutregion", C.c_int),
        ("nmesh", C.c_int), ("order", C.c_int), ("precision", C.c_int),
        ("force_scale", C.c_double),
    ]


class Pot(C.Structure):
    _fields_ = [
        ("npt", C.c_int), ("form", C.c_int), ("lewald", C.c_int), ("pad", C.c_int),
        ("alpha", C.c_double), ("rcut", C.c_double), ("epsi4pi", C.c_double),
        ("utoltab", C.c_double), ("ftoltab", C.c_double), ("umaxtab", C.c_double),
        ("fmaxtab", C.c_double), ("r2uminin", C.c_double),
        ("z", C.c_double * 8), ("sig", C.c_double * 8), ("eps", C.c_double * 8),
        ("rad", C.c_double * 8),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    """Load libmolsim_gpu.so (built by ``__graft_entry__.build()``); raise if absent."""
    global _lib
    if _lib is None:
        path = os.environ.get("MOLSIM_GPU_LIB", LIB_PATH)  # another build of the library, for A/B timing
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.msg_last_error.restype = C.c_char_p
        L.msg_kernel_name.restype = C.c_char_p
        L.msg_build_table.argtypes = [C.POINTER(Pot), _dp, C.c_int, _ip, _ip, _dp, _dp]
        L.msg_create.argtypes = [C.POINTER(Params), C.POINTER(_vp)]
        L.msg_destroy.argtypes = [_vp]
        L.msg_destroy.restype = None
        L.msg_set_table.argtypes = [_vp, _dp, C.c_int, _ip, _ip, _dp, _dp]
        L.msg_set_particles.argtypes = [_vp, _dp, _ip]
        L.msg_set_positions.argtypes = [_vp, _dp]
        L.msg_set_positions_device.argtypes = [_vp, _vp]
        L.msg_ewald_setup.argtypes = [_vp]
        L.msg_set_box.argtypes = [_vp, _dp]
        L.msg_nlist.argtypes = [_vp]
        L.msg_utotal.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.msg_utotal_async.argtypes = [_vp, _vp]
        L.msg_fetch.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.msg_dutotal.argtypes = [_vp, C.c_int, _ip, _dp, C.POINTER(C.c_int), _vp, _vp, _vp]
        L.msg_dutotal_batch.argtypes = [_vp, C.c_int, _ip, _dp, _ip, _dp, _vp]
        L.msg_dutotal_batch_noalloc.argtypes = [_vp, C.c_int, _ip, _dp, _ip, _dp, _vp]
        L.msg_commit.argtypes = [_vp, C.c_int, _ip, _dp]
        L.msg_set_weak_charges.argtypes = [_vp, _ip, _dp]
        L.msg_dutotal_charge.argtypes = [_vp, C.c_int, _ip, C.POINTER(C.c_int), _vp, _vp, _vp]
        L.msg_commit_charge.argtypes = [_vp, C.c_int, _ip]
        L.msg_get_cell_list.argtypes = [_vp, _ip, _vp, _vp]
        L.msg_md_init.argtypes = [_vp, _dp, _vp]
        L.msg_md_run.argtypes = [_vp, C.c_int, C.c_double, _vp]
        L.msg_get_positions.argtypes = [_vp, _dp]
        L.msg_get_velocities.argtypes = [_vp, _dp]
        L.msg_random.argtypes = [C.POINTER(C.c_int), _ip, C.POINTER(C.c_double)]
        L.msg_random.restype = C.c_double
        L.msg_mc_pass.argtypes = [_vp, C.c_int, C.c_int, _dp, C.c_double, C.POINTER(C.c_int), _ip,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), _vp]
        L.msg_nlist_builds.argtypes = [_vp]
        L.msg_nlist_builds.restype = C.c_longlong
        L.msg_nlist_interval.argtypes = [_vp, C.c_int]
        L.msg_nlist_macrostep.argtypes = [_vp]
        L.msg_set_nlist_width.argtypes = [_vp, C.c_int]
        L.msg_get_verlet_list.argtypes = [_vp, C.c_int, C.c_int, _ip, _ip]
        L.msg_stream.argtypes = [_vp]
        L.msg_stream.restype = _vp
        L.msg_profile.argtypes = [_vp, C.c_int]
        L.msg_profile_read.argtypes = [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.msg_device_bytes.argtypes = [_vp]
        L.msg_device_bytes.restype = C.c_longlong
        L.msg_set_tile.argtypes = [_vp, C.c_int]
        L.msg_get_tile.argtypes = [_vp]
        L.msg_get_forces.argtypes = [_vp, _dp]
        L.msg_query.argtypes = [_vp]
        L.msg_batch_create.argtypes = [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]
        L.msg_batch_destroy.argtypes = [_vp]
        L.msg_batch_destroy.restype = None
        L.msg_batch_utotal.argtypes = [_vp]
        L.msg_batch_sync.argtypes = [_vp]
        L.msg_batch_dutotal.argtypes = [_vp, _ip, _ip, _dp, _ip, _dp]
        L.msg_batch_mc_pass.argtypes = [_vp, C.c_int, C.c_int, _dp, C.c_double, _ip, _ip, _dp, _dp, _ip, _vp]
        L.msg_batch_results.argtypes = [_vp, _vp, _vp, _vp]
        L.msg_unit_create.argtypes = [C.POINTER(_vp)]
        L.msg_unit_destroy.argtypes = [_vp]
        L.msg_unit_destroy.restype = None
        L.msg_unit_launch.argtypes = [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, _vp]
        L.msg_unit_query.argtypes = [_vp]
        L.msg_unit_results.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.msg_batch_query.argtypes = [_vp]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise MolsimGpuError(rc, lib().msg_last_error().decode())
    return rc


def make_pot(pot: dict) -> Pot:
    p = Pot()
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


def build_table(pot: dict, maxbuf: int = 200000) -> dict:
    """PotTwoBodyTab: returns MOLSIM's table state (iubuflow 1-based, as MOLSIM holds it)."""
    p = make_pot(pot)
    npt = pot["npt"]
    nptpt = npt * (npt + 1) // 2
    ubuf = np.zeros(maxbuf)
    low = np.zeros(nptpt, dtype=np.int32)
    ng = np.zeros(nptpt, dtype=np.int32)
    r2umin = np.zeros(nptpt)
    r2atat = np.zeros(nptpt)
    n = _check(lib().msg_build_table(C.byref(p), ubuf, maxbuf, low, ng, r2umin, r2atat))
    return dict(ubuf=ubuf[:n].copy(), iubuflow=low, nugrid=ng, r2umin=r2umin, r2atat=r2atat)


class Random:
    """MOLSIM's Random(iseed) generator (Src/mollib.F90:2689-2713), state carried here."""

    def __init__(self, iseed: int):
        self.iseed = C.c_int(iseed)
        self.state = np.array([-1, -1], dtype=np.int32)
        self.am = C.c_double(0.0)

    def __call__(self) -> float:
        return lib().msg_random(C.byref(self.iseed), self.state, C.byref(self.am))

    @classmethod
    def from_state(cls, iseed: int, am: float, ix: int, iy: int) -> "Random":
        """Resume from a saved state (iseed, am, ix, iy), as a .cnf file carries it
        (Src/molsim.F90:876,936)."""
        g = cls(iseed)
        g.am.value = am
        g.state[:] = (ix, iy)
        return g

    def get_state(self) -> tuple[int, float, int, int]:
        """(iseed, am, ix, iy), the values IOCnf writes."""
        return self.iseed.value, self.am.value, int(self.state[0]), int(self.state[1])


def _ptr(a):
    return a.ctypes.data_as(_vp)


class Engine:
    """One device-resident system (MOLSIM's global state for the non-bonded path)."""

    def __init__(self, spec, precision: str = "mixed", table: dict | None = None,
                 force_scale: float = 1.0):
        self.spec = spec
        P = Params()
        P.np = spec.n
        P.npt = spec.npt
        for d in range(3):
            P.boxlen[d] = float(spec.box[d])
        P.rcut = spec.rcut
        P.drnlist = spec.drnlist
        P.epsi4pi = spec.pot["epsi4pi"]
        P.list_mode = LIST_MODES[spec.list_mode]
        P.ewald_mode = EWALD_MODES[spec.ewald]
        P.lsurf = int(spec.lsurf)
        P.alpha = spec.alpha
        P.ncut = spec.ncut
        P.ncutregion = {"sphere": 0, "cube": 1}[getattr(spec, "ncutregion", "sphere")]
        P.nmesh = spec.nmesh
        P.order = spec.order
        P.precision = PRECISIONS[precision]
        P.force_scale = force_scale
        self.params = P
        self.nptpt = spec.npt * (spec.npt + 1) // 2
        h = _vp()
        _check(lib().msg_create(C.byref(P), C.byref(h)))
        self.h = h
        self.table = table if table is not None else build_table(spec.pot)
        t = self.table
        _check(lib().msg_set_table(h, t["ubuf"], t["ubuf"].size, t["iubuflow"], t["nugrid"],
                                   t["r2umin"], t["r2atat"]))
        az = np.ascontiguousarray(spec.az, dtype=np.float64)
        ipt = np.ascontiguousarray(spec.ipt + 1, dtype=np.int32)
        _check(lib().msg_set_particles(h, az, ipt))
        self.set_positions(spec.ro)
        if getattr(spec, "laz", None) is not None:
            self.set_weak_charges(spec.laz, spec.pot["z"][: spec.npt])
        if spec.ewald != "none":
            self.ewald_setup()

    def close(self):
        if getattr(self, "h", None):
            lib().msg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------------------
    def set_positions(self, ro):
        ro = np.ascontiguousarray(ro, dtype=np.float64).reshape(-1)
        _check(lib().msg_set_positions(self.h, ro))

    def set_positions_device(self, ptr: int):
        _check(lib().msg_set_positions_device(self.h, _vp(ptr)))

    def ewald_setup(self):
        _check(lib().msg_ewald_setup(self.h))

    def set_box(self, boxlen):
        """Box change (SetBoxParam + EwaldSetup + InitCellList, Src/mc.F90:3423-3436); set the
        scaled positions afterwards."""
        box = np.ascontiguousarray(boxlen, dtype=np.float64).reshape(3)
        _check(lib().msg_set_box(self.h, box))
        self.spec = self.spec.__class__(**{**self.spec.__dict__, "box": box.copy()})

    def nlist(self):
        _check(lib().msg_nlist(self.h))

    # -- UTotal ---------------------------------------------------------------------------
    def utotal(self, want_force: bool = True) -> dict:
        ut = np.zeros(self.nptpt + 1)
        urec, utot, vir = C.c_double(), C.c_double(), C.c_double()
        f = np.zeros(3 * self.spec.n) if want_force else None
        _check(lib().msg_utotal(self.h, _ptr(ut), C.cast(C.pointer(urec), _vp),
                                C.cast(C.pointer(utot), _vp), _ptr(f) if f is not None else None,
                                C.cast(C.pointer(vir), _vp)))
        return dict(u_twob=ut, u_rec=urec.value, u_tot=utot.value,
                    force=None if f is None else f.reshape(-1, 3), virial=vir.value)

    def utotal_async(self, d_force: int | None = None):
        _check(lib().msg_utotal_async(self.h, _vp(d_force) if d_force else None))

    def fetch(self) -> dict:
        ut = np.zeros(self.nptpt + 1)
        urec, utot, vir = C.c_double(), C.c_double(), C.c_double()
        _check(lib().msg_fetch(self.h, _ptr(ut), C.cast(C.pointer(urec), _vp),
                               C.cast(C.pointer(utot), _vp), C.cast(C.pointer(vir), _vp)))
        return dict(u_twob=ut, u_rec=urec.value, u_tot=utot.value, virial=vir.value)

    # -- DUTotal / MCUpdate ---------------------------------------------------------------
    def dutotal(self, ip: int, rnew) -> dict:
        """Energy change of moving particle ip (0-based here) to rnew."""
        ipa = np.array([ip + 1], dtype=np.int32)
        r = np.ascontiguousarray(rnew, dtype=np.float64).reshape(3)
        ov = C.c_int()
        dt = np.zeros(self.nptpt + 1)
        drec, dtot = C.c_double(), C.c_double()
        _check(lib().msg_dutotal(self.h, 1, ipa, r, C.byref(ov), _ptr(dt),
                                 C.cast(C.pointer(drec), _vp), C.cast(C.pointer(dtot), _vp)))
        return dict(overlap=bool(ov.value), du_twob=dt, du_rec=drec.value, du_tot=dtot.value)

    def dutotal_batch(self, ips, rnew) -> dict:
        """Independent single-particle trials against the same state (0-based ips)."""
        ips = np.ascontiguousarray(np.asarray(ips) + 1, dtype=np.int32)
        r = np.ascontiguousarray(rnew, dtype=np.float64).reshape(-1)
        n = ips.size
        ov = np.zeros(n, dtype=np.int32)
        dt = np.zeros(n)
        dr = np.zeros(n)
        _check(lib().msg_dutotal_batch(self.h, n, ips, r, ov, dt, _ptr(dr)))
        return dict(overlap=ov.astype(bool), du_tot=dt, du_rec=dr)

    def dutotal_batch_noalloc(self, ips, rnew) -> dict:
        """dutotal_batch through the library's fixed scratch (no device allocation; at most
        16384 trials)."""
        ips = np.ascontiguousarray(np.asarray(ips) + 1, dtype=np.int32)
        r = np.ascontiguousarray(rnew, dtype=np.float64).reshape(-1)
        n = ips.size
        ov = np.zeros(n, dtype=np.int32)
        dt = np.zeros(n)
        dr = np.zeros(n)
        _check(lib().msg_dutotal_batch_noalloc(self.h, n, ips, r, ov, dt, _ptr(dr)))
        return dict(overlap=ov.astype(bool), du_tot=dt, du_rec=dr)

    def mc_update(self, ip: int, rnew):
        ipa = np.array([ip + 1], dtype=np.int32)
        _check(lib().msg_commit(self.h, 1, ipa, np.ascontiguousarray(rnew, dtype=np.float64).reshape(3)))

    # -- weak (titrating) charges --------------------------------------------------------------
    def set_weak_charges(self, laz, zat):
        """Charge states laz (N,) and full charges per type zat (npt,) (lweakcharge)."""
        _check(lib().msg_set_weak_charges(self.h, np.ascontiguousarray(laz, dtype=np.int32),
                                          np.ascontiguousarray(zat, dtype=np.float64)))

    def dutotal_charge(self, ips) -> dict:
        """ChargeChange trial: the charge states of particles ips (one or two, 0-based) flip."""
        ia = np.ascontiguousarray(np.atleast_1d(ips), dtype=np.int32) + 1
        dt = np.zeros(self.nptpt + 1)
        ov = C.c_int()
        dr, du = C.c_double(), C.c_double()
        _check(lib().msg_dutotal_charge(self.h, int(ia.size), ia, C.byref(ov), _ptr(dt), C.byref(dr),
                                        C.byref(du)))
        return dict(overlap=bool(ov.value), du_twob=dt, du_rec=dr.value, du_tot=du.value)

    def commit_charge(self, ips):
        """Accept a ChargeChange (MCUpdate of the charge states)."""
        ia = np.ascontiguousarray(np.atleast_1d(ips), dtype=np.int32) + 1
        _check(lib().msg_commit_charge(self.h, int(ia.size), ia))

    def cell_list(self):
        """(nc, cell_start, cell_ip 0-based) of the current positions (as of the last pass)."""
        nc = np.zeros(3, dtype=np.int32)
        _check(lib().msg_get_cell_list(self.h, nc, None, None))
        start = np.zeros(int(np.prod(nc)) + 1, dtype=np.int32)
        ips = np.zeros(self.spec.n, dtype=np.int32)
        _check(lib().msg_get_cell_list(self.h, nc, _ptr(start), _ptr(ips)))
        return nc, start, ips - 1

    def mc_pass(self, ntrial: int, dtran, beta: float, rng: "Random", isamp: int = 0):
        """MCPass of single-particle moves; returns (sum of accepted dE, naccept, events)."""
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dtran, dtype=np.float64),
                                                  (self.spec.npt,)))
        ev = np.zeros(ntrial, dtype=np.int32)
        dsum, nacc = C.c_double(), C.c_int()
        _check(lib().msg_mc_pass(self.h, int(ntrial), int(isamp), dt, float(beta), C.byref(rng.iseed),
                                 rng.state, C.byref(rng.am), C.byref(dsum), C.byref(nacc), _ptr(ev)))
        return dsum.value, nacc.value, ev

    # -- VelVer MD --------------------------------------------------------------------------
    def md_init(self, mass, vel=None):
        """Masses per particle type and initial velocities (N, 3) (None = zero)."""
        m = np.ascontiguousarray(mass, dtype=np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        _check(lib().msg_md_init(self.h, m, _ptr(v) if v is not None else None))

    def md_run(self, nsteps: int, dt: float, log: bool = True):
        out = np.zeros(2 * nsteps) if log else None
        _check(lib().msg_md_run(self.h, int(nsteps), float(dt), _ptr(out) if out is not None else None))
        return None if out is None else out.reshape(nsteps, 2)

    def positions(self):
        ro = np.zeros(3 * self.spec.n)
        _check(lib().msg_get_positions(self.h, ro))
        return ro.reshape(-1, 3)

    def velocities(self):
        v = np.zeros(3 * self.spec.n)
        _check(lib().msg_get_velocities(self.h, v))
        return v.reshape(-1, 3)

    def nlist_builds(self) -> int:
        return int(lib().msg_nlist_builds(self.h))

    def nlist_interval(self, inlist: int):
        """inlist: 0 rebuilds by the skin test, > 0 every inlist-th MD step of a macrostep."""
        _check(lib().msg_nlist_interval(self.h, int(inlist)))

    def set_nlist_width(self, maxnneigh: int):
        """The list width (MOLSIM's maxnneigh); 0 sizes it from the density."""
        _check(lib().msg_set_nlist_width(self.h, int(maxnneigh)))

    def nlist_macrostep(self):
        """NList(iBeforeMacrostep): displacement sums and the step count start again."""
        _check(lib().msg_nlist_macrostep(self.h))

    def verlet_list(self, kind: str = "md", maxnneigh: int = 256):
        """The Verlet list in MOLSIM's layout ("md": SetVListMD's parity half list, "mc":
        SetVListMC's full list): (nneighpn (N,), jpnlist (N, maxnneigh) of 0-based partners in
        ascending order, -1 past each row's end)."""
        n = self.spec.n
        nn = np.zeros(n, dtype=np.int32)
        jl = np.zeros(n * maxnneigh, dtype=np.int32)
        _check(lib().msg_get_verlet_list(self.h, {"md": 0, "mc": 1}[kind], int(maxnneigh), nn, jl))
        jl = jl.reshape(n, maxnneigh) - 1
        jl[np.arange(maxnneigh)[None, :] >= nn[:, None]] = -1
        return nn, jl

    # -- measurement ----------------------------------------------------------------------
    def stream(self) -> int:
        return lib().msg_stream(self.h) or 0

    def profile(self, enable: bool):
        _check(lib().msg_profile(self.h, int(enable)))

    def profile_read(self) -> dict:
        out = {}
        for k, name in enumerate(KERNELS):
            ms, n = C.c_double(), C.c_int()
            _check(lib().msg_profile_read(self.h, k, C.byref(ms), C.byref(n)))
            if n.value:
                out[name] = (ms.value, n.value)
        return out

    def device_bytes(self) -> int:
        return int(lib().msg_device_bytes(self.h))

    # -- tiles, completion, device-resident results --------------------------------------
    def set_tile(self, pair_lanes: int):
        _check(lib().msg_set_tile(self.h, int(pair_lanes)))

    def tile(self) -> int:
        return int(lib().msg_get_tile(self.h))

    def query(self) -> bool:
        rc = lib().msg_query(self.h)
        _check(min(rc, 0))
        return bool(rc)

    def forces(self):
        f = np.zeros(3 * self.spec.n)
        _check(lib().msg_get_forces(self.h, f))
        return f.reshape(-1, 3)


class Batch:
    """Resident systems evaluated together: each stage of the UTotal pass is one launch over
    all systems (captured once into a graph, so a batch step is one graph launch), and the
    MC trials of all systems are one launch."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (_vp * len(self.engines))(*[e.h for e in self.engines])
        h = _vp()
        _check(lib().msg_batch_create(arr, len(self.engines), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().msg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def utotal_async(self):
        _check(lib().msg_batch_utotal(self.h))

    def sync(self):
        _check(lib().msg_batch_sync(self.h))

    def query(self) -> bool:
        rc = lib().msg_batch_query(self.h)
        _check(min(rc, 0))
        return bool(rc)

    def energies(self) -> dict:
        """u_tot, u_rec, virial of every system (arrays), one download after the pass."""
        n = len(self.engines)
        ut, ur, vi = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().msg_batch_results(self.h, _ptr(ut), _ptr(ur), _ptr(vi)))
        return dict(u_tot=ut, u_rec=ur, virial=vi)

    def utotal(self, want_force: bool = True) -> list[dict]:
        self.utotal_async()
        self.sync()
        out = []
        for e in self.engines:
            r = e.fetch()
            r["force"] = e.forces() if want_force else None
            out.append(r)
        return out

    def dutotal(self, ips, rnews) -> list[dict]:
        """ips[s] (0-based indices) and rnews[s] (ntrial, 3) per system s."""
        nt = np.array([len(i) for i in ips], dtype=np.int32)
        ip = np.ascontiguousarray(np.concatenate([np.asarray(i) for i in ips]) + 1, dtype=np.int32)
        r = np.ascontiguousarray(np.concatenate([np.asarray(x).reshape(-1, 3) for x in rnews]),
                                 dtype=np.float64).reshape(-1)
        ovl = np.zeros(ip.size, dtype=np.int32)
        du = np.zeros(ip.size)
        _check(lib().msg_batch_dutotal(self.h, nt, ip, r, ovl, du))
        out, off = [], 0
        for n in nt:
            out.append(dict(overlap=ovl[off:off + n].astype(bool), du_tot=du[off:off + n]))
            off += n
        return out

    def mc_pass(self, ntrial: int, dtran, beta: float, rngs, isamp: int = 0, events: bool = True):
        """One MCPass per system, the chains side by side; rngs[s] is system s's Random
        (updated).  dtran: one value, or per system an array over its particle types.
        Returns (du_sum per system, naccept per system, events (nsys, ntrial) or None)."""
        n = len(self.engines)
        dt = np.zeros((n, MAX_TYPES))
        for s, e in enumerate(self.engines):
            d = np.asarray(dtran[s] if np.ndim(dtran) > 0 else dtran, dtype=np.float64)
            dt[s, :e.spec.npt] = np.broadcast_to(d, (e.spec.npt,))
        iseed = np.array([g.iseed.value for g in rngs], dtype=np.int32)
        st = np.ascontiguousarray(np.concatenate([g.state for g in rngs]), dtype=np.int32)
        am = np.array([g.am.value for g in rngs])
        dsum, nacc = np.zeros(n), np.zeros(n, dtype=np.int32)
        ev = np.zeros((n, ntrial), dtype=np.int32) if events else None
        _check(lib().msg_batch_mc_pass(self.h, int(ntrial), int(isamp), np.ascontiguousarray(dt), float(beta),
                                       iseed, st, am, dsum, nacc, _ptr(ev) if ev is not None else None))
        for s, g in enumerate(rngs):
            g.iseed.value = int(iseed[s])
            g.state[:] = st[2 * s:2 * s + 2]
            g.am.value = float(am[s])
        return dsum, nacc, ev


class Unit:
    """Execution unit of the request loop: one batch of requests at a time (a UTotal pass over
    some resident systems, MC trials of others) as segmented launches on its own streams."""

    def __init__(self):
        h = _vp()
        _check(lib().msg_unit_create(C.byref(h)))
        self.h = h
        self._keep = None
        self._shape = None

    def close(self):
        if getattr(self, "h", None):
            lib().msg_unit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def launch(self, ef_engines, de_engines=(), trials=(), want_forces: bool = True):
        """ef_engines: systems for energy + forces; de_engines with trials[s] = (ips 0-based,
        rnew (n, 3)).  Returns at once."""
        ef = list(ef_engines)
        de = list(de_engines)
        ua = (_vp * max(1, len(ef)))(*[e.h for e in ef])
        da = (_vp * max(1, len(de)))(*[e.h for e in de])
        nt = np.array([len(t[0]) for t in trials], dtype=np.int32) if de else np.zeros(1, np.int32)
        if de:
            ip = np.ascontiguousarray(np.concatenate([np.asarray(t[0]) for t in trials]) + 1, dtype=np.int32)
            r = np.ascontiguousarray(np.concatenate([np.asarray(t[1]).reshape(-1, 3) for t in trials]),
                                     dtype=np.float64).reshape(-1)
        else:
            ip = np.zeros(1, np.int32)
            r = np.zeros(3)
        _check(lib().msg_unit_launch(self.h, ua, len(ef), int(want_forces), da, len(de),
                                     _ptr(nt), _ptr(ip), _ptr(r)))
        self._keep = (ua, da, nt, ip, r)
        self._shape = (ef, de, nt[:len(de)].copy(), want_forces)

    def done(self) -> bool:
        rc = lib().msg_unit_query(self.h)
        _check(min(rc, 0))
        return bool(rc)

    def results(self):
        """(per ef system: dict(u_tot, u_rec, virial, force)), (per de system: dict(overlap, du_tot))."""
        ef, de, nt, wf = self._shape
        n = len(ef)
        ut, ur, vi = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
        nf = sum(3 * e.spec.n for e in ef)
        f = np.zeros(max(nf, 1)) if wf else None
        ntot = int(nt.sum())
        ovl = np.zeros(max(ntot, 1), dtype=np.int32)
        du = np.zeros(max(ntot, 1))
        _check(lib().msg_unit_results(self.h, _ptr(ut), _ptr(ur), _ptr(vi),
                                      _ptr(f) if f is not None else None, _ptr(ovl), _ptr(du)))
        out_ef, off = [], 0
        for i, e in enumerate(ef):
            d = dict(u_tot=ut[i], u_rec=ur[i], virial=vi[i], force=None)
            if f is not None:
                d["force"] = f[off:off + 3 * e.spec.n].reshape(-1, 3).copy()
                off += 3 * e.spec.n
            out_ef.append(d)
        out_de, off = [], 0
        for k in nt:
            out_de.append(dict(overlap=ovl[off:off + k].astype(bool), du_tot=du[off:off + k].copy()))
            off += k
        return out_ef, out_de
