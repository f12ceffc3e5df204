"""The oracle's per-cell Z* (species of different Z*, read_relaxor.F:35) pinned
against its one-Z* path.

dipoFFT.F:35-44 transforms Z*(R) u and dVddi-dipole-dipole.F:40-62 applies
L(k) to it and multiplies the result by Z*(R), so the long-range energy of u
with per-cell Z*(R) is the one-Z* energy (Z0) of v(R) = Z*(R) u(R) / Z0, and
the long-range force is Z*(R) / Z0 times the one-Z* force of v.  The
short-range energy does not see Z*; the field energy is -sum Z*(R) u.E
(mirror_force_mirror.F:39-65).  A dropped Z*(R) on either side of the
convolution, or Z0 in place of Z*(R), breaks these identities.
"""
import numpy as np

from feram_amd.params import batio3_zvr

ZS = [9.956, 7.5, 12.25]
E = [0.0007, -0.0011, 0.0023]


def _oracle(oracle_mod, p, u, zr=None, relax=None):
    o = oracle_mod.Oracle(p)
    if relax is not None:
        o.relaxor[...] = relax
    if zr is not None:
        o.field("Z_starR")[...] = zr
    o.field("E_field")[...] = np.array(p.external_E_field)
    o.field("dipoR")[...] = u
    e = o.mirror_force_mirror().as_dict()
    out = {"e": e, "F_long": o.field("dVddi_long").copy(), "F": o.field("dVddi").copy()}
    o.close()
    return out


def test_species_zstar_reduces_to_one_zstar(oracle_mod):
    p = batio3_zvr((8, 6, 10))
    p.Z_star = list(ZS)
    p.external_E_field = list(E)
    p.E_field_always_zero = False
    rng = np.random.default_rng(44)
    relax = rng.integers(1, 4, size=(p.Lz, p.Ly, p.Lx)).astype(np.int32)
    zr = np.array(ZS)[relax - 1]
    u = 0.07 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    a = _oracle(oracle_mod, p, u, zr, relax)

    q = batio3_zvr((8, 6, 10))
    q.Z_star = [ZS[0]] * 3
    v = u * (zr / ZS[0])[..., None]
    b = _oracle(oracle_mod, q, v)
    assert abs(a["e"]["long_range"] - b["e"]["long_range"]) <= 1e-12 * abs(b["e"]["long_range"])
    ratio = (zr / ZS[0])[..., None]
    assert np.max(np.abs(a["F_long"] - ratio * b["F_long"])) <= 1e-12 * np.max(np.abs(b["F_long"]))
    # the field term with the cell's Z*
    ede = -np.sum(zr[..., None] * u * np.array(E))
    assert abs(a["e"]["dipole_E_field"] - ede) <= 1e-12 * abs(ede)
    # one Z* per species equal to Z0 is the one-Z* system
    c = _oracle(oracle_mod, q, u, np.full(zr.shape, ZS[0]), relax)
    d = _oracle(oracle_mod, q, u)
    assert c["e"]["long_range"] == d["e"]["long_range"]


def test_species_zstar_force_is_the_energy_gradient(oracle_mod):
    p = batio3_zvr((6, 6, 8))
    p.Z_star = list(ZS)
    rng = np.random.default_rng(45)
    relax = rng.integers(1, 4, size=(p.Lz, p.Ly, p.Lx)).astype(np.int32)
    zr = np.array(ZS)[relax - 1]
    u = 0.06 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    a = _oracle(oracle_mod, p, u, zr, relax)
    terms = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain", "homo_coupling",
             "inho_strain", "inho_coupling")
    h = 1e-5
    for (z, y, x, k) in [(0, 0, 0, 0), (3, 2, 1, 2), (7, 5, 4, 1), (5, 1, 3, 0)]:
        up, um = u.copy(), u.copy()
        up[z, y, x, k] += h
        um[z, y, x, k] -= h
        Ep = sum(_oracle(oracle_mod, p, up, zr, relax)["e"][t] for t in terms)
        Em = sum(_oracle(oracle_mod, p, um, zr, relax)["e"][t] for t in terms)
        fd = (Ep - Em) / (2 * h)
        assert abs(fd - a["F"][z, y, x, k]) <= 1e-6 * max(1.0, abs(fd)), (z, y, x, k, fd, a["F"][z, y, x, k])
