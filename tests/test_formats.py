"""The .coord / .restart format (feram_amd.formats) against a Fortran compiler.

write_system's record is the Fortran format (i4,2i5,3f10.6,9e14.6,3f13.6)
(coord_module.F:474-479).  A small Fortran program compiled here with
amdflang writes records of the same values with that format; the Python
formatter must produce the same text byte for byte (values chosen to cover
signs, zeros, rounding carries into the next decade and tiny and large
exponents).  Parsing the text back gives the values to the printed precision.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from feram_amd import formats

FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

VALUES = [0.0, 1.0, -1.0, 0.1, 0.9999996, 0.99999949, 1.234567e-5, -9.8765432e-12, 123456.789, -0.5e-99,
          3.3e100, 0.035, -0.1000004, 0.09999999, 5e-7, -2.5e-8, 0.0001, 7.0, -3.14159265358979, 6.02e23]


def records(rng, n):
    return rng.choice(VALUES, size=(n, 15)) * rng.choice([1.0, 1.0, 0.5, -1.0], size=(n, 15))


def test_e_descriptor_known_answers():
    assert formats.fortran_e(1.0) == "  0.100000E+01"
    assert formats.fortran_e(-0.00123456) == " -0.123456E-02"
    assert formats.fortran_e(0.9999996) == "  0.100000E+01"
    assert formats.fortran_e(0.0) == "  0.000000E+00"
    assert formats.fortran_e(1.5e-101) == "  0.150000-100"


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler")
def test_coord_records_match_fortran(tmp_path):
    rng = np.random.default_rng(0)
    recs = records(rng, 40)
    recs[:, :3] = np.clip(recs[:, :3], -99.0, 99.0)        # f10.6 fields hold displacements
    recs[:, 12:] = np.clip(recs[:, 12:], -99999.0, 99999.0)  # f13.6
    np.savetxt(tmp_path / "in.txt", recs, fmt="%.17e")
    (tmp_path / "w.f90").write_text("""
program w
  implicit none
  real*8 :: v(15)
  integer :: i, ios
  open(10, file='in.txt', status='old')
  do i = 0, 39
     read(10, *, iostat=ios) v
     if (ios /= 0) exit
     write(*, '(i4,2i5,3f10.6,9e14.6,3f13.6)') mod(i, 4), mod(i / 4, 5), i / 20, v
  end do
end program w
""")
    subprocess.run([FLANG, "w.f90", "-o", "w"], cwd=tmp_path, check=True)
    ref = subprocess.run(["./w"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    a = recs.reshape(2, 5, 4, 15)
    text = formats.format_coord(a[..., 0:3], a[..., 3:6], a[..., 6:9], a[..., 9:12], a[..., 12:15])
    assert text.splitlines() == ref


def test_parse_round_trip():
    rng = np.random.default_rng(1)
    shp = (3, 4, 5, 3)
    f = {k: rng.normal(0, s, shp) for k, s in (("dipoR", 0.1), ("dipoP", 0.3), ("dVddi", 2.0),
                                               ("acouR", 0.01), ("acouP", 0.02))}
    text = formats.format_coord(f["dipoR"], f["dipoP"], f["dVddi"], f["acouR"], f["acouP"])
    g = formats.parse_restart(text, 5, 4, 3)
    assert np.max(np.abs(g["dipoR"] - f["dipoR"])) <= 5e-7
    assert np.max(np.abs(g["acouP"] - f["acouP"])) <= 5e-7
    for k in ("dipoP", "dVddi", "acouR"):
        assert np.all(np.abs(g[k] - f[k]) <= 5e-6 * np.abs(f[k]) + 1e-300)
    with pytest.raises(ValueError):
        formats.parse_restart(text, 4, 5, 3)
