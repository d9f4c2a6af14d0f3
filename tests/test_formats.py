"""CPU: the reference's output and restart formats (suguru_amd.formats).

Pins: Ew.d editing equals what a Fortran compiler writes for the same values (golden fixture
from flang, tests/golden/make_fortran_e_golden.py); the binary restart files round-trip bit
for bit and are read back by a Fortran program that declares the reference's own record
layout (the particle_restart derived type and a direct-access record, loadd.f90:151-188; the
interior arrays of loadflds, loadd.f90:84-103), compiled here with flang.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from suguru_amd import formats as F

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fortran_e_format.npz")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def test_e_editing_matches_fortran():
    z = np.load(GOLD)
    v = z["values"]
    assert F.format_lines(v.reshape(-1, 3).T, 15, 7) == str(z["e157"])
    assert F.format_lines(v.reshape(-1, 7).T, 16, 8) == str(z["e168"])


def test_negative_zero_without_sign_and_specials():
    assert F.fortran_e(-0.0, 15, 7) == "  0.0000000E+00"
    assert F.fortran_e(float("nan"), 15, 7).strip() == "NaN"
    assert F.fortran_e(-1e300, 15, 7) == " -0.1000000+301"


def test_capsule_file_name_and_columns(tmp_path):
    assert F.capsule_txt_name(7, 1200) == "Capsule007step0001200.txt"
    x, y, z = np.array([1.5, 2.25]), np.array([0.125, -3.0]), np.array([10.0, 0.0])
    p = tmp_path / F.capsule_txt_name(1, 100)
    F.write_capsule_txt(str(p), x, y, z)
    lines = p.read_text().splitlines()
    assert lines[0] == "  0.1250000E+00  0.1000000E+02  0.1500000E+01"
    assert lines[1] == " -0.3000000E+01  0.0000000E+00  0.2250000E+01"
    rx, ry, rz = F.read_capsule_txt(str(p))
    assert np.array_equal(rx, x) and np.array_equal(ry, y) and np.array_equal(rz, z)


def test_center_velocity_and_timestep_lines(tmp_path):
    p = tmp_path / "center_velocity.txt"
    F.append_center_velocity(str(p), 0.01, (1.0, 2.0, 3.0), (0.5, -0.25, 0.0), lines=2)
    t = p.read_text().splitlines()
    assert len(t) == 2 and t[0] == t[1]
    assert t[0] == ("  0.10000000E-01  0.10000000E+01  0.20000000E+01  0.30000000E+01"
                    "  0.50000000E+00 -0.25000000E+00  0.00000000E+00")
    q = tmp_path / "timestep.txt"
    F.append_timestep(str(q), 100, 0.01, 1e-4)
    assert q.read_text() == "  0.10000000E+03  0.10000000E-01  0.10000000E-03\n"


def _fields(itot, jtot, ktot, seed):
    rng = np.random.default_rng(seed)
    return {n: np.asfortranarray(rng.standard_normal((itot, jtot, ktot))) for n in F.FLD_ORDER}


def test_fldbg_round_trip_and_layout(tmp_path):
    itot, jtot, ktot = 6, 4, 5
    f = _fields(itot, jtot, ktot, 1)
    I, J, K = np.meshgrid(np.arange(itot), np.arange(jtot), np.arange(ktot), indexing="ij")
    f["vnew"] = np.asfortranarray(I + 100.0 * J + 10000.0 * K)
    p = str(tmp_path / F.fldbg_name(300))
    F.write_fldbg(p, f, 0.03, 300, 1e-4)
    raw = np.fromfile(p, dtype="<f8")
    n = itot * jtot * ktot
    assert raw.size == 7 * n + 3
    # vnew is the second array; element (i, j, k) at i + itot (j + jtot k)
    assert raw[n + 2 + itot * (3 + jtot * 4)] == 2 + 100.0 * 3 + 10000.0 * 4
    assert list(raw[-3:]) == [0.03, 300.0, 1e-4]
    g, time, nr, dt = F.read_fldbg(p, itot, jtot, ktot)
    assert (time, nr, dt) == (0.03, 300, 1e-4)
    for k in F.FLD_ORDER:
        assert np.array_equal(g[k], f[k])


def _records(npart, nl, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(npart, dtype=F.particle_record_dtype(nl))
    for name in F.PARTICLE_SCALARS:
        r[name] = rng.standard_normal(npart)
    for name in F.PARTICLE_ARRAYS:
        r[name] = rng.standard_normal((npart, nl))
    return r


def test_allpartdata_round_trip_and_layout(tmp_path):
    npart, nl = 3, 10
    r = _records(npart, nl, 2)
    p = str(tmp_path / F.allpartdata_name(100))
    F.write_allpartdata(p, r)
    assert os.path.getsize(p) == npart * (6 + 22 * nl) * 8
    raw = np.fromfile(p, dtype="<f8").reshape(npart, 6 + 22 * nl)
    assert np.array_equal(raw[:, 3], r["integralv"])
    k = F.PARTICLE_ARRAYS.index("fzl")
    assert np.array_equal(raw[:, 6 + k * nl:6 + (k + 1) * nl], r["fzl"])
    back = F.read_allpartdata(p, npart, nl)
    assert back.tobytes() == r.tobytes()


FORTRAN_READER = """program rd
  implicit none
  integer, parameter :: np = {npart}, nl = {nl}, itot = {itot}, jtot = {jtot}, ktot = {ktot}
  type particle_restart
    real(8) :: x, y, z, integralv, volume, volume0
    real(8), dimension(nl) :: xfp, yfp, zfp, dxdt, dydt, dzdt, fxl, fyl, fzl, ul, vl, wl, &
                              xfpo, yfpo, zfpo, xfpold, yfpold, zfpold, xfp0, yfp0, zfp0, area
  end type particle_restart
  type(particle_restart), dimension(np) :: glob
  real(8) :: temp(itot, jtot, ktot), info(3)
  integer :: k
  open(20, file='part.bin', access='direct', form='unformatted', recl=np*(6+22*nl)*8)
  read(20, rec=1) glob
  close(20)
  write(*, '(4ES26.17)') glob(2)%volume0, glob(3)%yfp(nl), glob(1)%area(1), glob(2)%xfpold(4)
  open(21, file='fld.bin', access='stream', form='unformatted')
  do k = 1, 6
    read(21) temp
  end do
  write(*, '(2ES26.17)') temp(2, 3, 4), temp(itot, 1, ktot)
  read(21) temp
  read(21) info
  close(21)
  write(*, '(3ES26.17)') info
end program rd
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not in this image")
def test_fortran_reads_the_files_with_the_reference_layout(tmp_path):
    npart, nl, itot, jtot, ktot = 3, 10, 6, 4, 5
    r = _records(npart, nl, 5)
    F.write_allpartdata(str(tmp_path / "part.bin"), r)
    f = _fields(itot, jtot, ktot, 6)
    F.write_fldbg(str(tmp_path / "fld.bin"), f, 1.25, 7, 5e-5)
    (tmp_path / "rd.f90").write_text(FORTRAN_READER.format(npart=npart, nl=nl, itot=itot,
                                                           jtot=jtot, ktot=ktot))
    subprocess.run([FLANG, "-O0", "rd.f90", "-o", "rd"], cwd=tmp_path, check=True)
    out = subprocess.run(["./rd"], cwd=tmp_path, check=True, capture_output=True,
                         text=True).stdout.split()
    got = [float(s) for s in out]
    want = [r["volume0"][1], r["yfp"][2, nl - 1], r["area"][0, 0], r["xfpold"][1, 3],
            f["wo"][1, 2, 3], f["wo"][itot - 1, 0, ktot - 1], 1.25, 7.0, 5e-5]
    assert got == want
