"""The reference's .coord / .restart text format for device state.

write_system (coord_module.F:434-486) writes one line per cell, z slowest and
x fastest, with the Fortran format (i4,2i5,3f10.6,9e14.6,3f13.6):
ix iy iz, dipoR(3), dipoP(3) / s_Nose, dVddi(3), acouR(3), acouP(3); with
optimised acoustic displacements (mass_acou <= 0) acouR is the w that
coord_module.F:457-469 reconstructs from acouK.  initialize_dipoR
(initialize-dipoR.F:22-37) reads the same 15 numbers back from a .restart
file (list-directed) into dipoR, dipoP, dVddi, acouR, acouP; the following
mirror_force_mirror recomputes dVddi, and s_Nose restarts at 1.
"""
from __future__ import annotations

import math

import numpy as np


def fortran_e(v: float, w: int = 14, d: int = 6) -> str:
    """Fortran Ew.d edit descriptor (0.ddddddE+xx, exponent digits as
    gfortran prints them), right-aligned in w characters."""
    v = float(v)
    if v == 0.0 or not math.isfinite(v):
        if v == 0.0:
            s = ("-" if math.copysign(1.0, v) < 0 else "") + f"{0.0:.{d}f}E+00"
        else:
            s = "NaN" if math.isnan(v) else ("-" if v < 0 else "") + "Infinity"
        return s.rjust(w)
    sign = "-" if v < 0 else ""
    # d significant digits, correctly rounded from the binary value
    mant, ex = f"{abs(v):.{d - 1}e}".split("e")
    digits = "." + mant.replace(".", "")
    e = int(ex) + 1
    exp = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return (sign + "0" + digits + exp).rjust(w)


def format_coord(dipoR, dipoP, dVddi, acouR, acouP) -> str:
    """The .coord text of arrays shaped (Lz, Ly, Lx, 3); dipoP already divided by s_Nose."""
    Lz, Ly, Lx, _ = np.shape(dipoR)
    out = []
    for iz in range(Lz):
        for iy in range(Ly):
            for ix in range(Lx):
                parts = [f"{ix:4d}{iy:5d}{iz:5d}"]
                parts += [f"{v:10.6f}" for v in dipoR[iz, iy, ix]]
                parts += [fortran_e(v) for v in dipoP[iz, iy, ix]]
                parts += [fortran_e(v) for v in dVddi[iz, iy, ix]]
                parts += [fortran_e(v) for v in acouR[iz, iy, ix]]
                parts += [f"{v:13.6f}" for v in acouP[iz, iy, ix]]
                out.append("".join(parts))
    return "\n".join(out) + "\n"


def parse_restart(text: str, Lx: int, Ly: int, Lz: int) -> dict:
    """Fields of a .restart / .coord text (list-directed read, one cell a line)."""
    vals = np.array(text.split(), dtype=float)
    n = Lx * Ly * Lz
    if vals.size != 18 * n:
        raise ValueError(f"expected {n} lines of 18 numbers, got {vals.size} numbers")
    rows = vals.reshape(n, 18)
    out = {}
    for k, name in enumerate(("dipoR", "dipoP", "dVddi", "acouR", "acouP")):
        out[name] = rows[:, 3 + 3 * k: 6 + 3 * k].reshape(Lz, Ly, Lx, 3)
    idx = rows[:, :3].astype(int)
    expect = np.stack(np.meshgrid(np.arange(Lx), np.arange(Ly), np.arange(Lz), indexing="ij"), axis=-1)
    if not np.array_equal(idx.reshape(Lz, Ly, Lx, 3), expect.transpose(2, 1, 0, 3)):
        raise ValueError("cell indices out of order (z slowest, x fastest)")
    return out


def write_coord(system, path: str) -> None:
    """write_system for a device context."""
    p = system.p
    s = system.nose[0]
    acoustic = p.mass_acou[0] > 0.0
    zeros = np.zeros((p.Lz, p.Ly, p.Lx, 3))
    text = format_coord(system.get("dipoR"), system.get("dipoP") / s, system.get("dVddi"),
                        system.get("acouR") if acoustic else system.get("acouR_optimised"),
                        system.get("acouP") if acoustic else zeros)
    with open(path, "w") as f:
        f.write(text)


def read_restart(system, path: str) -> dict:
    """initialize_dipoR's .restart branch: put u, p (and w, p_w with acoustic
    MD) on the device; the caller then runs md_prologue (force, s = 1, pi = 0,
    H0) as initialize_dipoR and molecular_dynamics do."""
    p = system.p
    with open(path) as f:
        fields = parse_restart(f.read(), p.Lx, p.Ly, p.Lz)
    system.put("dipoR", fields["dipoR"])
    system.put("dipoP", fields["dipoP"])
    if p.mass_acou[0] > 0.0:
        system.put("acouR", fields["acouR"])
        system.put("acouP", fields["acouP"])
    return fields


# ---------------------------------------------------------------------------
# .avg: one record per run (average_module.F:180-260)
# ---------------------------------------------------------------------------

# The record's edit descriptors in order: kelvin f8.3, E(3) f9.5, strain(6),
# <u>(3), <uu>(6) as (e15.5, e13.5...), 15 energies as (e17.7, e15.7...),
# <p>(3), <pp>(6) as (e15.5, e13.5...) (average_module.F:240).
_AVG_GROUPS = [(1, "f", 8, 3), (3, "f", 9, 5), (6, "e", 13, 5), (3, "e", 13, 5), (6, "e", 13, 5),
               (15, "e", 15, 7), (3, "e", 13, 5), (6, "e", 13, 5)]
_AVG_ENERGIES = ("dipo_kinetic", "long_range", "dipole_E_field", "unharmonic", "homo_strain", "homo_coupling",
                 "inho_strain", "inho_coupling", "total", "H_Nose_Poincare")


def format_avg_values(values) -> str:
    """The .avg record of its 43 values, in the reference's format
    (f8.3,3f9.5,1(e15.5,5e13.5),1(e15.5,2e13.5),1(e15.5,5e13.5),1(1e17.7,14e15.7),
    1(e15.5,2e13.5),1(e15.5,5e13.5)); the first E field of each e group is
    two characters wider."""
    v = [float(x) for x in values]
    if len(v) != 43:
        raise ValueError(f"an .avg record holds 43 values, got {len(v)}")
    out, k = [], 0
    for n, kind, w, d in _AVG_GROUPS:
        for j in range(n):
            x = v[k]
            k += 1
            if kind == "f":
                s = f"{x:{w}.{d}f}"
                out.append(s if len(s) <= w else "*" * w)
            else:
                out.append(fortran_e(x, w + 2 if j == 0 else w, d))
    return "".join(out)


def avg_values(avg: dict, p, n_sum_p: int | None = None, n_average: int | None = None) -> list:
    """The 43 values of divide_and_write_Average (average_module.F:198-259)
    from the Average_type sums of System.avg_get().

    n_average defaults to the number of averaging steps summed; n_sum_p is
    the count of sum_p cells (film and epit divide <u>, <uu>, <p>, <pp> by
    it instead of N).  For 'lf' the written temperature is the one the run
    reached (average_module.F:218-224); the other methods write p.kelvin."""
    from .params import KELVIN_EV
    n = int(avg["n_add"] if n_average is None else n_average)
    if n <= 0:
        raise ValueError("no averaging step has been summed (n_average = 0 writes no record)")
    N = p.N
    Ninv = 1.0 / N
    e = {k: v / n for k, v in avg["energy"].items()}                 # div_Energy(avg%energy, n_average)
    e = {k: v / N for k, v in e.items()}                              # div_Energy(avg%energy, N)
    e2 = avg["e2"] / n * Ninv
    dkt = avg["dipo_kinetic_true"] / n * Ninv
    strain = [s / n for s in avg["strain"]]
    if p.bulk_or_film in ("film", "epit"):
        if not n_sum_p:
            raise ValueError("film and epit records need n_sum_p")
        norm = lambda xs: [x / n / n_sum_p for x in xs]                 # noqa: E731
    else:
        norm = lambda xs: [x / n * Ninv for x in xs]                    # noqa: E731
    u, uu, pm, pp = norm(avg["u"]), norm(avg["uu"]), norm(avg["p"]), norm(avg["pp"])
    kelvin = p.kelvin
    if p.method == "lf":
        if p.mass_acou[0] > 0.0:
            kelvin = (dkt + e["acou_kinetic"]) / 3 * KELVIN_EV
        else:
            kelvin = dkt * 2 / 3 * KELVIN_EV
    vals = [kelvin] + list(p.external_E_field[:3]) + strain + u + uu
    vals += [e[k] for k in _AVG_ENERGIES] + [e2, dkt, e["acou_kinetic"], e["short_range"], e["inho_modulation"]]
    vals += pm + pp
    return vals


def write_avg(system, path: str, n_average: int | None = None) -> str:
    """Write the .avg record of a System's device-side averages to path
    (replacing the file, as the reference opens it with status='REPLACE')
    and return the line."""
    avg = system.avg_get()
    n_sum_p = int(system.counted().sum()) if system.film else None
    line = format_avg_values(avg_values(avg, system.p, n_sum_p, n_average))
    with open(path, "w") as f:
        f.write(line + "\n")
    return line
