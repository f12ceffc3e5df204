"""The oracle against what the mathematics fixes.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * short-range kernel eigenvalues at Gamma, X, M, R equal the closed forms the
    reference prints (dipole-dipole-short-range.F:102-113);
  * the Ewald long-range kernel is independent of the splitting parameter
    kappa, and at Gamma equals the Lorentz value -(4 pi/3)/a0^3 * 14.399644/eps;
  * u = 0 under hydrostatic pressure gives eta = -P a0^3/(B11+2B12)/160.22
    (GPa_check.sh:83);
  * uniform u along 001/110/111 gives the polynomials of
    potential_surface_check.sh:96-116;
  * optimised strains satisfy E_coupling = -2 E_strain (md24x24x24a.r1567.log:106-109).
"""
import math

import numpy as np
import pytest

from conftest import load_golden, param_from_dict
from feram_amd.params import batio3_zvr


def test_short_range_special_points(oracle_mod):
    p = batio3_zvr(8)
    o = oracle_mod.Oracle(p)
    S = o.field("short_range")
    k2, j = p.P_kappa2[0], p.j
    j1, j2, j3, j4, j5, j6, j7 = j
    gamma = k2 + 2 * j1 + j2 + 4 * j3 + 2 * j4 + 4 * j6
    x_lo = k2 + 2 * j1 - j2 - 4 * j3 + 2 * j4 - 4 * j6
    x_to = k2 + j2 - 2 * j4 - 4 * j6
    m_to = k2 - 2 * j1 + j2 - 4 * j3 + 2 * j4 + 4 * j6
    m_lo = k2 - j2 - 2 * j4 + 4 * j6
    r = k2 - 2 * j1 - j2 + 4 * j3 + 2 * j4 - 4 * j6
    h = 4
    cases = {(0, 0, 0): [gamma] * 3, (h, 0, 0): [x_lo, x_to, x_to],
             (h, h, 0): [m_to, m_lo, m_lo], (h, h, h): [r] * 3}
    for (ix, iy, iz), want in cases.items():
        ev = np.linalg.eigvalsh(S[iz, iy, ix].T) / 2
        assert np.allclose(np.sort(ev), np.sort(want), rtol=0, atol=1e-13)


def test_ewald_kappa_independence_and_gamma(oracle_mod):
    p = batio3_zvr(8)
    o = oracle_mod.Oracle(p)
    base = o.field("long_range").copy()
    kap = p.kappa_ewald
    for f in (0.75, 1.35):
        o2 = oracle_mod.Oracle(p.copy(kappa=f * kap))
        other = o2.field("long_range")
        assert np.max(np.abs(other - base)) <= 1e-10 * np.max(np.abs(base))
    lorentz = -(4 * math.pi / 3) / p.a0 ** 3 * 14.399644 / p.epsilon_inf
    assert np.allclose(base[0, 0, 0], lorentz * np.eye(3), rtol=0, atol=1e-12 * abs(lorentz))
    # symmetric in (a, b) and even in k: L(k) = L(-k) on the grid
    assert np.allclose(base, np.swapaxes(base, -1, -2), rtol=0, atol=1e-15 * np.max(np.abs(base)))
    Lz, Ly, Kx = base.shape[:3]
    neg = base[(-np.arange(Lz)) % Lz][:, (-np.arange(Ly)) % Ly]
    assert np.allclose(neg[:, :, 0], base[:, :, 0], rtol=0, atol=1e-12 * np.max(np.abs(base)))


@pytest.mark.parametrize("gpa", [-5.0, 3.0])
def test_pressure_strain_closed_form(oracle_mod, gpa):
    p = batio3_zvr(6)
    p.GPa = [gpa] * 3
    o = oracle_mod.Oracle(p)
    o.mirror_force_mirror()
    want = -gpa * p.a0 ** 3 / (p.B11 + 2 * p.B12) / (100 * 1.60217653)
    assert np.allclose(o.strain[:3], want, rtol=1e-13, atol=0)
    assert np.allclose(o.strain[3:], 0.0, atol=1e-18)


def test_potential_surface_polynomials(oracle_mod):
    g = load_golden()["dispersion_check"]
    p = param_from_dict(g["params"], Lx=4, Ly=4, Lz=4, GPa=[0.0, 0.0, 0.0])
    o = oracle_mod.Oracle(p)
    kappa_short = p.P_kappa2[0] + 2 * p.j[0] + p.j[1] + 4 * p.j[2] + 2 * p.j[3] + 4 * p.j[5]
    kappa = kappa_short - 14.399644 * 2.0 / 3.0 * math.pi * p.Z_star[0] ** 2 / p.epsilon_inf / p.a0 ** 3
    B = p.B11 + 2 * p.B12
    Cc = p.B1xx + 2 * p.B1yy
    mu_t, nu_t = (p.B11 - p.B12) / 2, (p.B1xx - p.B1yy) / 2
    mu_r, nu_r = p.B44, p.B4yz
    ap = p.P_alpha[0] - (Cc ** 2 / B + 4 * nu_t ** 2 / mu_t) / 24.0
    gp = p.P_gamma[0] + (nu_t ** 2 / mu_t - nu_r ** 2 / mu_r) / 2.0
    k1, k2, k3, k4 = p.P_k1[0], p.P_k2[0], p.P_k3[0], p.P_k4[0]
    forms = {
        (0, 0, 1): lambda x: kappa * x ** 2 + ap * x ** 4 + k1 * x ** 6 + k4 * x ** 8,
        (1, 1, 0): lambda x: kappa * x ** 2 + (ap + gp / 4) * x ** 4 + (k1 + k2 / 4) * x ** 6 + k4 * x ** 8,
        (1, 1, 1): lambda x: kappa * x ** 2 + (ap + gp / 3) * x ** 4
        + (k1 + 2 * k2 / 9 + k3 / 27) * x ** 6 + k4 * x ** 8,
    }
    for d, E in forms.items():
        d = np.array(d, float) / np.linalg.norm(d)
        for x in (0.05, 0.17, 0.31):
            o.field("dipoR")[...] = x * d
            e = o.mirror_force_mirror()
            got = oracle_mod.lib().fo_sum_potential_energies(e) / p.N
            assert got == pytest.approx(E(x), rel=1e-11, abs=1e-14)


def test_optimised_strain_energy_identities(oracle_mod):
    p = batio3_zvr(8)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(5)
    o.field("dipoR")[...] = 0.1 + 0.05 * rng.standard_normal((8, 8, 8, 3))
    e = o.mirror_force_mirror()
    assert e.homo_coupling == pytest.approx(-2 * e.homo_strain, rel=1e-12)
    assert e.inho_coupling == pytest.approx(-2 * e.inho_strain, rel=1e-11)
    assert e.inho_strain > 0.0
