"""MOLSIM configuration files (molsim_amd.cnf): the record layout of IOCnf
(Src/molsim.F90:839-948) against a Fortran program compiled here with the same I/O lists
(tests/cnf_io.F90, amdflang), round trips, error handling, and the request loop's ingest and
emit of system states.  The reference's own configuration file is checked in
test_oracle_ref_mollib.py."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from molsim_amd import cnf, systems

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fortran(tmp_path_factory):
    if not shutil.which("amdflang"):
        pytest.skip("amdflang not in the image")
    d = tmp_path_factory.mktemp("cnf_io")
    exe = str(d / "cnf_io")
    subprocess.run(["amdflang", "-O0", os.path.join(HERE, "cnf_io.F90"), "-o", exe], check=True,
                   cwd=str(d), capture_output=True)
    return exe


def _fortran_values(np_=5):
    """The values tests/cnf_io.F90 writes."""
    ip = np.arange(1, np_ + 1)[:, None].astype(float)
    d = np.arange(1, 4)[None, :].astype(float)
    ro = ip + 0.125 * d - 3.0
    rod = -ip - 0.5 * d
    qua = np.tile([0.5, -0.5, 0.5, -0.5], (np_, 1))
    quad = np.column_stack([np.zeros(np_), ip[:, 0], np.full(np_, -2.0), np.full(np_, 0.25)])
    laz = (np.arange(1, np_ + 1) % 2) == 0
    return ro, qua, rod, quad, laz


def test_fortran_written_file_reads(fortran, tmp_path):
    path = tmp_path / "f.cnf"
    subprocess.run([fortran, "write", str(path)], check=True)
    c = cnf.read_cnf(path, lweakcharge=True, lmd=True)
    ro, qua, rod, quad, laz = _fortran_values()
    assert (c.istep, c.iseed, c.ix, c.iy) == (42, 4712, -993160733, 1260042462)
    assert c.am == 4.6566125955216364e-10
    assert np.array_equal(c.boxlen, [10.5, 11.25, 12.125])
    assert np.array_equal(c.ro, ro) and np.array_equal(c.qua, qua)
    assert np.array_equal(c.rod, rod) and np.array_equal(c.quad, quad)
    assert np.array_equal(c.laz, laz)


def test_python_written_file_reads_in_fortran(fortran, tmp_path):
    rng = np.random.default_rng(3)
    n = 5
    c = cnf.Cnf(istep=123456, boxlen=np.array([80.0, 81.5, 79.25]), ro=rng.uniform(-40, 40, (n, 3)),
                qua=rng.normal(size=(n, 4)), iseed=99, am=4.6566125955216364e-10, ix=-7, iy=123,
                laz=rng.integers(0, 2, n).astype(bool), rod=rng.normal(size=(n, 3)), quad=rng.normal(size=(n, 4)))
    path = tmp_path / "p.cnf"
    cnf.write_cnf(path, c, lweakcharge=True, lmd=True)
    out = subprocess.run([fortran, "read", str(path)], check=True, capture_output=True, text=True).stdout.split("\n")
    head = out[0].split()
    assert (int(head[0]), int(head[1]), int(head[3]), int(head[4])) == (c.istep, c.iseed, c.ix, c.iy)
    assert float(head[2]) == c.am
    assert np.array_equal(np.array(out[1].split(), dtype=float), c.boxlen)
    assert int(out[2]) == n
    for i in range(n):
        f = out[3 + i].split()
        v = np.array(f[:14], dtype=float)
        assert np.array_equal(v[:3], c.ro[i]) and np.array_equal(v[3:7], c.qua[i])
        assert np.array_equal(v[7:10], c.rod[i]) and np.array_equal(v[10:14], c.quad[i])
        assert (f[14] == "T") == bool(c.laz[i])


def test_round_trip_all_records(tmp_path):
    rng = np.random.default_rng(1)
    n, m = 37, 3
    c = cnf.Cnf(istep=7, boxlen=np.array([50.0, 50.0, 50.0]), ro=rng.uniform(-25, 25, (n, 3)),
                qua=rng.normal(size=(n, 4)), iseed=4712, am=0.5, ix=1, iy=2,
                nbondcl=rng.integers(0, m + 1, n).astype(np.int32),
                bondcl=rng.integers(0, n, (n, m)).astype(np.int32), laz=rng.integers(0, 2, n).astype(bool),
                rod=rng.normal(size=(n, 3)), quad=rng.normal(size=(n, 4)))
    path = tmp_path / "r.cnf"
    cnf.write_cnf(path, c, lclink=True, lweakcharge=True, lmd=True)
    d = cnf.read_cnf(path, lclink=True, maxvalnbondcl=m, lweakcharge=True, lmd=True)
    for k in ("istep", "iseed", "am", "ix", "iy"):
        assert getattr(d, k) == getattr(c, k)
    for k in ("boxlen", "ro", "qua", "nbondcl", "bondcl", "laz", "rod", "quad"):
        assert np.array_equal(getattr(d, k), getattr(c, k)), k
    # 'zero' start keeps step and box, not the generator state (Src/molsim.F90:864-874)
    z = cnf.read_cnf(path, txstart="zero", lntp=True)
    assert z.istep == 7 and np.array_equal(z.boxlen, c.boxlen) and z.iseed == 0 and z.ix == -1
    # lmd without a velocity record: tolerated, as the reference's end = 998
    cnf.write_cnf(path, c)
    assert cnf.read_cnf(path, lmd=True).rod is None


def test_malformed_files_raise(tmp_path):
    c = cnf.Cnf(istep=1, boxlen=np.ones(3), ro=np.zeros((4, 3)))
    path = tmp_path / "m.cnf"
    cnf.write_cnf(path, c)
    raw = path.read_bytes()
    bad = tmp_path / "bad.cnf"
    bad.write_bytes(raw[:-3])
    with pytest.raises(cnf.CnfError):
        cnf.read_cnf(bad)
    bad.write_bytes(raw[:-4] + (999).to_bytes(4, "little"))
    with pytest.raises(cnf.CnfError, match="trailing marker"):
        cnf.read_cnf(bad)
    with pytest.raises(cnf.CnfError, match="needed"):
        cnf.read_cnf(path, np_expected=5)
    with pytest.raises(cnf.CnfError, match="laz record missing"):
        cnf.read_cnf(path, lweakcharge=True)
    with pytest.raises(cnf.CnfError):
        cnf.read_cnf(path, txstart="restart")


def test_apply_to_spec_wraps_and_checks_count():
    sp = systems.small_spme_system(n=50, L=30.0, nmesh=16, alpha=0.3, rcut=10.0)
    ro = sp.ro + np.array([30.0, -30.0, 60.0])         # periodic images
    c = cnf.Cnf(istep=0, boxlen=np.zeros(3), ro=ro)      # box not in the file: kept
    s2 = cnf.apply_to_spec(sp, c)
    assert np.allclose(s2.ro, sp.ro, atol=1e-12) and np.array_equal(s2.box, sp.box)
    assert np.all(s2.ro >= -s2.box / 2) and np.all(s2.ro < s2.box / 2)
    with pytest.raises(cnf.CnfError):
        cnf.apply_to_spec(sp, cnf.Cnf(istep=0, boxlen=np.zeros(3), ro=ro[:10]))


def test_server_ingests_and_emits(tmp_path):
    from molsim_amd.serve import Server

    class Eng:
        def __init__(self, sp):
            self.spec, self.ro, self.closed = sp, None, False

        def set_positions(self, ro):
            self.ro = np.array(ro)

        def close(self):
            self.closed = True

    sp = systems.small_spme_system(n=40, L=30.0, nmesh=16, alpha=0.3, rcut=10.0)
    srv = Server({0: sp}, table={}, hbm_budget=1 << 30, units=0, engine_factory=Eng,
                 footprint_fn=lambda s: 1000)
    eng = srv._admit_system(0)
    moved = np.array(sp.ro)
    moved[3] *= 0.5                                      # stays inside the box
    path = tmp_path / "in.cnf"
    cnf.write_cnf(path, cnf.Cnf(istep=5, boxlen=sp.box, ro=moved))
    srv.ingest_cnf(0, path)
    assert np.allclose(eng.ro, moved) and srv.resident[0] is eng
    out = tmp_path / "out.cnf"
    srv.emit_cnf(0, out, istep=6)
    back = cnf.read_cnf(out)
    assert back.istep == 6 and np.allclose(back.ro, moved)
    # a new box evicts the resident context; the next admission builds from the new spec
    cnf.write_cnf(path, cnf.Cnf(istep=5, boxlen=sp.box * 1.1, ro=moved))
    srv.ingest_cnf(0, path)
    assert eng.closed and 0 not in srv.resident and srv.bytes == 0
    assert np.allclose(srv._admit_system(0).spec.box, sp.box * 1.1)
