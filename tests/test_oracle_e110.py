"""The oracle's 'e110' homogeneous strain (optimize-homo-strain.F:42-56) pinned
against the constrained minimum of the homogeneous energy.

e110 fixes eta_3 = f (epi_strain), eta_4 = eta_5 = 0 and takes
eta = (f + d, f + d, f, 0, 0, -2 d) with d chosen by a closed form.  The
homogeneous energy is E(eta) = eta.H.eta / 2 + eta.c (energy of
optimize-homo-strain.F:77-85) with c the coupling vector; the bulk solve
eta_bulk = -H^-1 c (pinned in test_oracle_closed_forms) gives c = -H eta_bulk
for the same u.  Minimising E over d gives
    d* = [-(H11 + H22 + 2 H21) f - (H31 + H32) f - c1 - c2 + 2 c6] / (H11 + H22 + 2 H21 + 4 H66).
The reference's numerator has 2*homogeneous(2,2) where the minimum has
2*homogeneous(2,1) (optimize-homo-strain.F:47); its denominator has H21.
The oracle keeps the reference's formula, so:
  * at f = 0 the slip vanishes and the oracle's eta is the constrained minimum
    (a brute-force scan of E(d) checks the closed form too);
  * at f != 0 the oracle's d differs from d* by exactly
    -2 (H22 - H21) f / (H11 + H22 + 2 H21 + 4 H66).
A "corrected" oracle (H21 in the numerator) fails the second check, any other
change of sign or index fails the first.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr


def _state(oracle_mod, geometry, f, seed=5):
    p = batio3_zvr((8, 8, 8), method="md", kelvin=300.0)
    p.GPa = [-4.0, -6.0, -2.0]
    p.bulk_or_film = geometry
    p.epi_strain = f
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(seed)
    u = 0.08 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("dipoR")[...] = u
    e = o.mirror_force_mirror().as_dict()
    H = np.ctypeslib.as_array(oracle_mod.lib().fo_homogeneous(o.ctx), shape=(36,)).reshape(6, 6).copy()
    eta = o.strain.copy()
    o.close()
    return H, eta, e


def _eta(f, d):
    return np.array([f + d, f + d, f, 0.0, 0.0, -2.0 * d])


def _coupling(oracle_mod):
    H, eta_bulk, _ = _state(oracle_mod, "bulk", 0.0)
    return H, -H @ eta_bulk


def _minimum(H, c, f):
    den = H[0, 0] + H[1, 1] + 2 * H[1, 0] + 4 * H[5, 5]
    num = -(H[0, 0] + H[1, 1] + 2 * H[1, 0]) * f - (H[2, 0] + H[2, 1]) * f - c[0] - c[1] + 2 * c[5]
    return num / den, den


def test_e110_closed_form_is_the_constrained_minimum(oracle_mod):
    H, c = _coupling(oracle_mod)
    f = 0.004
    d_star, _ = _minimum(H, c, f)
    energy = lambda d: 0.5 * _eta(f, d) @ H @ _eta(f, d) + _eta(f, d) @ c
    ds = d_star + np.linspace(-1e-3, 1e-3, 2001)
    es = np.array([energy(d) for d in ds])
    assert abs(ds[np.argmin(es)] - d_star) <= 1.1e-6
    # stationary: dE/dd = 0 at d*
    h = 1e-7
    assert abs(energy(d_star + h) - energy(d_star - h)) <= 1e-9 * abs(energy(d_star))


def test_e110_oracle_at_zero_epitaxial_strain_is_the_minimum(oracle_mod):
    H, c = _coupling(oracle_mod)
    _, eta, e = _state(oracle_mod, "e110", 0.0)
    d_star, _ = _minimum(H, c, 0.0)
    assert np.allclose(eta, _eta(0.0, d_star), rtol=1e-12, atol=1e-15)
    assert e["homo_strain"] == pytest.approx(0.5 * eta @ H @ eta, rel=1e-12)
    assert e["homo_coupling"] == pytest.approx(eta @ c, rel=1e-10)


def test_e110_oracle_keeps_the_reference_numerator(oracle_mod):
    H, c = _coupling(oracle_mod)
    f = 0.004
    _, eta, _ = _state(oracle_mod, "e110", f)
    d_star, den = _minimum(H, c, f)
    assert eta[2] == f and eta[3] == 0.0 and eta[4] == 0.0
    d = eta[0] - f
    assert eta[1] == eta[0] and eta[5] == pytest.approx(-2 * d, rel=1e-15)
    slip = -2 * (H[1, 1] - H[1, 0]) * f / den
    assert abs(slip) > 1e-4 * abs(d_star)
    assert d - d_star == pytest.approx(slip, rel=1e-9)
