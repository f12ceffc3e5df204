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
