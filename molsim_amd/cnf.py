"""MOLSIM configuration files (.cnf): reader and writer.

A .cnf file is a Fortran unformatted sequential file written by IOCnf
(Src/molsim.F90:839-948).  Each record is framed by a 4-byte little-endian length marker
before and after its payload.  The records, in order:

  1. istep1 int32, iseed int32, am real(8), ix int32, iy int32, boxlen(3) real(8)
     -- 48 bytes; written by IOCnf('write') (:936) and read by txstart = 'continue' (:876).
     Older files (2007-03-16 -- 2017-08-22) hold nstep1done int32, iaux int32, boxlen(3)
     real(8) -- 32 bytes (:868); txstart = 'zero' without NPT reads only nstep1done (:872).
  2. np int32, iaux int32, raux real(8), raux real(8), laux logical(4), laux logical(4)
     -- 32 bytes (:882, :937).  Only np carries information.
  3. (ro(1:3,ip), qua(0:3,ip), ip = 1,np) -- 7 real(8) per particle (:884, :938).
  4. if lclink: nbondcl(1:np) int32, bondcl(1:maxvalnbondcl,1:np) int32 (:886-894, :939).
  5. if lweakcharge: laz(1:np) logical(4) (:895-898, :940).
  6. if lmd: (rod(1:3,ip), quad(0:3,ip), ip = 1,np) -- 7 real(8) per particle (:899-903,
     :941); the reader tolerates its absence (end = 998).

(iseed, am, ix, iy) is the state of MOLSIM's generator Random (Src/mollib.F90:2684-2713):
``Random(iseed)`` in :mod:`molsim_amd.engine` carries the same four values, so a device
MC chain can continue a MOLSIM run where its .cnf left off and hand its state back.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

_MAX_RECORD = 2**31 - 1    # one 4-byte marker; larger records would need sub-records


@dataclass
class Cnf:
    """Contents of a .cnf file.  Arrays are (np, 3) / (np, 4), row i = particle i + 1,
    the memory layout of Fortran ro(3,np) / qua(0:3,np)."""

    istep: int
    boxlen: np.ndarray
    ro: np.ndarray
    qua: np.ndarray | None = None
    iseed: int = 0
    am: float = 0.0
    ix: int = -1
    iy: int = -1
    nbondcl: np.ndarray | None = None
    bondcl: np.ndarray | None = None      # (np, maxvalnbondcl)
    laz: np.ndarray | None = None         # (np,) bool
    rod: np.ndarray | None = None
    quad: np.ndarray | None = None

    @property
    def np(self) -> int:
        return int(self.ro.shape[0])

    def rng_state(self) -> tuple[int, float, int, int]:
        """(iseed, am, ix, iy) of MOLSIM's Random."""
        return self.iseed, self.am, self.ix, self.iy


class CnfError(ValueError):
    pass


def _records(buf: bytes):
    """Split a sequential unformatted file into record payloads (4-byte markers)."""
    out, pos = [], 0
    while pos < len(buf):
        if pos + 4 > len(buf):
            raise CnfError(f"truncated record marker at byte {pos}")
        (n,) = struct.unpack_from("<i", buf, pos)
        if n < 0:
            raise CnfError(f"record at byte {pos} is split into sub-records (> 2 GiB); not supported")
        end = pos + 4 + n
        if end + 4 > len(buf):
            raise CnfError(f"record at byte {pos} runs past the end of the file")
        (m,) = struct.unpack_from("<i", buf, end)
        if m != n:
            raise CnfError(f"record at byte {pos}: leading marker {n} != trailing marker {m}")
        out.append(buf[pos + 4:end])
        pos = end + 4
    return out


def _frame(payload: bytes) -> bytes:
    if len(payload) > _MAX_RECORD:
        raise CnfError(f"record of {len(payload)} bytes exceeds one 4-byte marker")
    m = struct.pack("<i", len(payload))
    return m + payload + m


def _logical(a) -> np.ndarray:
    """Fortran logical(4) values to bool (any nonzero is .true.)."""
    return np.frombuffer(a, dtype="<i4") != 0


def read_cnf(path, *, txstart: str = "continue", lntp: bool = False, lclink: bool = False,
             maxvalnbondcl: int = 0, lweakcharge: bool = False, lmd: bool = False,
             np_expected: int | None = None) -> Cnf:
    """Read a .cnf as IOCnf('read') does (Src/molsim.F90:860-904).

    The flags are those of the run that reads it (MOLSIM knows them from its input).
    Record 1 is decoded by its length: 48 bytes is the current layout (step, generator
    state, box), 32 bytes the 2007-03-16 -- 2017-08-22 one (step, an unused integer, box).
    txstart 'continue' needs the current layout (:876); 'zero' keeps the step count and
    the box only (:864-874) and leaves the generator state at its defaults.

    The particle count is ``np_expected`` when given: MOLSIM reads its own np particles
    and uses record 2's count only for lmvt (:883).  Otherwise it is record 2's count, or,
    where that is zero (as in older files), the size of record 3 / 56."""
    with open(path, "rb") as f:
        recs = _records(f.read())
    if len(recs) < 3:
        raise CnfError(f"{path}: {len(recs)} records, at least 3 expected")
    r1, r2, r3 = recs[0], recs[1], recs[2]
    if txstart not in ("zero", "continue"):
        raise CnfError(f"txstart {txstart!r}: 'zero' or 'continue'")
    c = Cnf(istep=0, boxlen=np.zeros(3), ro=np.zeros((0, 3)))
    if len(r1) >= 48:
        istep, iseed, am, ix, iy = struct.unpack_from("<iidii", r1, 0)
        c.boxlen = np.frombuffer(r1, dtype="<f8", count=3, offset=24).copy()
        c.istep = istep
        if txstart == "continue":
            c.iseed, c.am, c.ix, c.iy = iseed, am, ix, iy
    elif len(r1) >= 32:
        if txstart == "continue":
            raise CnfError(f"record 1 holds {len(r1)} bytes (2007-2017 layout, no generator "
                           "state); txstart 'continue' needs the 48-byte layout")
        (c.istep,) = struct.unpack_from("<i", r1, 0)
        c.boxlen = np.frombuffer(r1, dtype="<f8", count=3, offset=8).copy()
    elif len(r1) >= 4 and txstart == "zero" and not lntp:
        (c.istep,) = struct.unpack_from("<i", r1, 0)
    else:
        raise CnfError(f"record 1 holds {len(r1)} bytes; not a .cnf header")
    if len(r2) < 4:
        raise CnfError("record 2 holds no particle count")
    (npread,) = struct.unpack_from("<i", r2, 0)
    if np_expected is not None:
        npart = int(np_expected)
    elif npread > 0:
        npart = npread
    else:
        npart = len(r3) // 56
    if len(r3) < 56 * npart:
        raise CnfError(f"record 3 holds {len(r3)} bytes; {56 * npart} needed for {npart} particles")
    a = np.frombuffer(r3, dtype="<f8", count=7 * npart).reshape(npart, 7)
    c.ro = a[:, :3].copy()
    c.qua = a[:, 3:].copy()
    k = 3
    if lclink:
        if k >= len(recs):
            raise CnfError("lclink: bond record missing")
        v = np.frombuffer(recs[k], dtype="<i4")
        if v.size < npart * (1 + maxvalnbondcl):
            raise CnfError("lclink: bond record too short")
        c.nbondcl = v[:npart].copy()
        c.bondcl = v[npart:npart * (1 + maxvalnbondcl)].reshape(npart, maxvalnbondcl).copy()
        k += 1
    if lweakcharge:
        if k >= len(recs):
            raise CnfError("lweakcharge: laz record missing")
        if len(recs[k]) < 4 * npart:
            raise CnfError("lweakcharge: laz record too short")
        c.laz = _logical(recs[k][:4 * npart])
        k += 1
    if lmd and k < len(recs):
        if len(recs[k]) < 56 * npart:
            raise CnfError("lmd: velocity record too short")
        v = np.frombuffer(recs[k], dtype="<f8", count=7 * npart).reshape(npart, 7)
        c.rod = v[:, :3].copy()
        c.quad = v[:, 3:].copy()
    return c


def write_cnf(path, c: Cnf, *, lclink: bool = False, lweakcharge: bool = False,
              lmd: bool = False) -> None:
    """Write a .cnf as IOCnf('write') does (Src/molsim.F90:932-942).  Missing
    orientations are written as the identity quaternion (1, 0, 0, 0), as OriToQua gives
    for monoatomic particles' unit orientation; record 2's unused fields are zero."""
    npart = c.np
    ro = np.ascontiguousarray(c.ro, dtype="<f8").reshape(npart, 3)
    qua = np.tile([1.0, 0.0, 0.0, 0.0], (npart, 1)) if c.qua is None else c.qua
    out = [_frame(struct.pack("<iidii", int(c.istep), int(c.iseed), float(c.am), int(c.ix), int(c.iy))
                  + np.asarray(c.boxlen, dtype="<f8").reshape(3).tobytes()),
           _frame(struct.pack("<iiddii", npart, 0, 0.0, 0.0, 0, 0)),
           _frame(np.hstack([ro, np.asarray(qua, dtype="<f8").reshape(npart, 4)]).tobytes())]
    if lclink:
        if c.nbondcl is None or c.bondcl is None:
            raise CnfError("lclink: nbondcl and bondcl required")
        out.append(_frame(np.asarray(c.nbondcl, dtype="<i4").reshape(npart).tobytes()
                          + np.ascontiguousarray(c.bondcl, dtype="<i4").reshape(npart, -1).tobytes()))
    if lweakcharge:
        if c.laz is None:
            raise CnfError("lweakcharge: laz required")
        out.append(_frame(np.asarray(c.laz, dtype=bool).astype("<i4").tobytes()))
    if lmd:
        if c.rod is None:
            raise CnfError("lmd: rod required")
        quad = np.zeros((npart, 4)) if c.quad is None else c.quad
        out.append(_frame(np.hstack([np.asarray(c.rod, dtype="<f8").reshape(npart, 3),
                                     np.asarray(quad, dtype="<f8").reshape(npart, 4)]).tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))


# -- system state in and out ------------------------------------------------------------------

def apply_to_spec(spec, c: Cnf):
    """A copy of ``spec`` (molsim_amd.systems.SystemSpec) with the positions and, when the
    file holds one, the box of ``c``; positions are wrapped into [-L/2, L/2) as MOLSIM keeps
    them.  The particle count must match (the file carries no types or charges)."""
    import dataclasses
    if c.np != spec.n:
        raise CnfError(f"file holds {c.np} particles, the system {spec.n}")
    box = np.asarray(c.boxlen, dtype=np.float64) if np.all(np.asarray(c.boxlen) > 0) else spec.box
    ro = np.asarray(c.ro, dtype=np.float64)
    ro = ro - box * np.floor(ro / box + 0.5)
    ro = np.where(ro >= box / 2, ro - box, ro)
    return dataclasses.replace(spec, ro=np.ascontiguousarray(ro), box=np.array(box, dtype=np.float64))


def from_spec(spec, istep: int = 0, rng=None, rod=None) -> Cnf:
    """A Cnf of a system description; ``rng`` (molsim_amd.engine.Random) supplies the
    generator state, ``rod`` (N, 3) the velocities of an MD state."""
    iseed, am, ix, iy = rng.get_state() if rng is not None else (0, 0.0, -1, -1)
    return Cnf(istep=int(istep), boxlen=np.array(spec.box, dtype=np.float64), ro=np.array(spec.ro),
               iseed=iseed, am=am, ix=ix, iy=iy, rod=None if rod is None else np.asarray(rod))


def from_engine(eng, istep: int = 0, rng=None, velocities: bool = False) -> Cnf:
    """A Cnf of a device-resident system: its current positions (and velocities after
    msg_md_init), downloaded from the device."""
    c = from_spec(eng.spec, istep, rng)
    c.ro = eng.positions()
    if velocities:
        c.rod = eng.velocities()
    return c
