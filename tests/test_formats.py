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


# ---- .avg record (average_module.F:180-260) --------------------------------

AVG_FMT = "(f8.3,3f9.5,1(e15.5,5e13.5),1(e15.5,2e13.5),1(e15.5,5e13.5),1(1e17.7,14e15.7),1(e15.5,2e13.5),1(e15.5,5e13.5))"


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler")
def test_avg_record_matches_fortran(tmp_path):
    rng = np.random.default_rng(3)
    recs = rng.choice(VALUES, size=(12, 43)) * rng.choice([1.0, -1.0, 0.5], size=(12, 43))
    recs[:, 0] = np.abs(np.clip(recs[:, 0], -999.0, 999.0))          # kelvin, f8.3
    recs[:, 1:4] = np.clip(recs[:, 1:4], -9.0, 9.0)                 # E field, f9.5
    recs[3, :] = 0.0
    np.savetxt(tmp_path / "in.txt", recs, fmt="%.17e")
    (tmp_path / "a.f90").write_text(f"""
program a
  implicit none
  real*8 :: v(43)
  integer :: i
  open(10, file='in.txt', status='old')
  do i = 1, 12
     read(10, *) v
     write(*, "{AVG_FMT}") v
  end do
end program a
""")
    subprocess.run([FLANG, "a.f90", "-o", "a"], cwd=tmp_path, check=True)
    ref = subprocess.run(["./a"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    assert [formats.format_avg_values(r) for r in recs] == ref


def _avg_sums(n, per_cell, N_norm, n_cells):
    """Average_type sums of n averaging steps whose per-step values are per_cell."""
    e = {k: per_cell.get(k, 0.0) * n * n_cells for k in
         ("acou_kinetic", "dipo_kinetic", "short_range", "long_range", "dipole_E_field", "unharmonic",
          "homo_strain", "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation", "total",
          "H_Nose_Poincare")}
    return {"energy": e, "e2": per_cell.get("e2", 0.0) * n * n_cells,
            "dipo_kinetic_true": per_cell.get("dkt", 0.0) * n * n_cells,
            "strain": [0.01 * n] * 6, "u": [0.1 * n * N_norm, -0.2 * n * N_norm, 0.3 * n * N_norm],
            "uu": [0.04 * n * N_norm] * 6, "p": [0.0] * 3, "pp": [0.5 * n * N_norm] * 6, "n_add": n}


def test_avg_normalisation_pins():
    from feram_amd.params import KELVIN_EV, batio3_zvr
    T = 287.5
    kT = T / KELVIN_EV
    # bulk: sums over N cells and n steps come back as per-cell means; a
    # leapfrog run whose dipole kinetic energy is 3/2 kT per cell writes T
    p = batio3_zvr(8, method="lf", kelvin=0.0)
    a = _avg_sums(7, {"dkt": 1.5 * kT, "unharmonic": 0.25, "e2": 0.0625}, p.N, p.N)
    v = formats.avg_values(a, p)
    assert v[0] == pytest.approx(T, rel=1e-14)
    assert v[4:10] == pytest.approx([0.01] * 6, rel=1e-14)
    assert v[10:13] == pytest.approx([0.1, -0.2, 0.3], rel=1e-14)
    assert v[22] == pytest.approx(0.25, rel=1e-14)          # unharmonic
    assert v[29] == pytest.approx(0.0625, rel=1e-14)        # e2
    assert v[30] == pytest.approx(1.5 * kT, rel=1e-14)      # dipo_kinetic_true
    # with acoustic MD the two kinetic energies (3/2 kT each) are averaged
    pa = batio3_zvr(8, method="lf", kelvin=0.0, acoustic_mass_amu=46.44)
    a = _avg_sums(5, {"dkt": 1.5 * kT, "acou_kinetic": 1.5 * kT}, pa.N, pa.N)
    assert formats.avg_values(a, pa)[0] == pytest.approx(T, rel=1e-14)
    # other methods write the set temperature; film/epit divide <u> by n_sum_p
    pm = batio3_zvr(8, method="md", kelvin=300.0).copy(bulk_or_film="epit")
    n_sum_p = 8 * 8 * 6
    a = _avg_sums(3, {"dkt": 9.0}, n_sum_p, pm.N)
    v = formats.avg_values(a, pm, n_sum_p=n_sum_p)
    assert v[0] == 300.0
    assert v[10:13] == pytest.approx([0.1, -0.2, 0.3], rel=1e-14)
    line = formats.format_avg_values(v)
    assert len(line.split()) == 43 and line.startswith(" 300.000")
    with pytest.raises(ValueError):
        formats.avg_values(a, pm)                            # film without n_sum_p
    with pytest.raises(ValueError):
        formats.avg_values(dict(a, n_add=0), p)              # nothing summed
