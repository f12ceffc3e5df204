"""The oracle's forces and integrators against invariants.

  * Every force is the derivative of the reported potential energy: a central
    finite difference of V(u) (and of V(w) with acoustic MD) reproduces
    dVddi (and dVdacR).  With the strains optimised, dV/du at the optimum is
    the partial derivative (envelope theorem), so this also checks the strain
    couplings.
  * Leapfrog conserves the total energy with an O(dt^2) error
    (zzzlf.gp:20-39 plots it for inspection); halving dt divides it by ~4.
  * The Nose-Poincare Hamiltonian stays near zero (nose-poincare.F:6-7,
    molecular-dynamics.F:115 "H_Nose_Poincare == 0, always").
  * Velocity scaling holds the kinetic energy at 1.5 N kT (param_module.F:468).
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr


def _random_state(o, p, seed, with_w=False):
    rng = np.random.default_rng(seed)
    o.field("dipoR")[...] = 0.08 + 0.06 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    if with_w:
        o.field("acouR")[...] = 0.02 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))


def _V(o, oracle_mod):
    e = o.mirror_force_mirror()
    return oracle_mod.lib().fo_sum_potential_energies(e)


@pytest.mark.parametrize("efield", [False, True])
def test_force_is_gradient_of_energy(oracle_mod, efield):
    p = batio3_zvr((6, 4, 8))
    if efield:
        p.external_E_field = [0.002, -0.001, 0.003]
        p.E_field_always_zero = False
    o = oracle_mod.Oracle(p)
    _random_state(o, p, 11)
    o.mirror_force_mirror()
    F = o.field("dVddi").copy()
    u = o.field("dipoR")
    h = 1e-5
    rng = np.random.default_rng(3)
    for _ in range(6):
        idx = (rng.integers(p.Lz), rng.integers(p.Ly), rng.integers(p.Lx), rng.integers(3))
        u0 = u[idx]
        u[idx] = u0 + h
        vp = _V(o, oracle_mod)
        u[idx] = u0 - h
        vm = _V(o, oracle_mod)
        u[idx] = u0
        fd = (vp - vm) / (2 * h)
        assert fd == pytest.approx(F[idx], rel=2e-6, abs=2e-8 * np.max(np.abs(F)))


def test_acoustic_force_is_gradient_of_energy(oracle_mod):
    p = batio3_zvr((6, 6, 4), method="lf", acoustic_mass_amu=46.44)
    o = oracle_mod.Oracle(p)
    _random_state(o, p, 12, with_w=True)
    o.mirror_force_mirror()
    Fw = o.field("dVdacR").copy()
    Fu = o.field("dVddi").copy()
    w = o.field("acouR")
    u = o.field("dipoR")
    h = 1e-6
    rng = np.random.default_rng(4)
    for arr, F in ((w, Fw), (u, Fu)):
        for _ in range(4):
            idx = (rng.integers(p.Lz), rng.integers(p.Ly), rng.integers(p.Lx), rng.integers(3))
            a0 = arr[idx]
            arr[idx] = a0 + h
            vp = _V(o, oracle_mod)
            arr[idx] = a0 - h
            vm = _V(o, oracle_mod)
            arr[idx] = a0
            fd = (vp - vm) / (2 * h)
            assert fd == pytest.approx(F[idx], rel=1e-5, abs=1e-7 * np.max(np.abs(F)))


def _lf_energy_error(oracle_mod, dt, t_total):
    p = batio3_zvr(6, method="lf")
    p.dt = dt
    o = oracle_mod.Oracle(p)
    _random_state(o, p, 21)
    e = o.md_prologue()
    minv = 1.0 / p.mass_dipo[0]
    # p is staggered (p_{n-1/2} -> p_{n+1/2}); the on-step kinetic energy uses
    # their mean, which makes E conserved to O(dt^2).
    E0 = None
    err = 0.0
    for i in range(int(round(t_total / dt))):
        p_old = o.field("dipoP").copy()
        o.leapfrog(e, i + 1)
        pm = 0.5 * (p_old + o.field("dipoP"))
        E = 0.5 * minv * np.sum(pm * pm) + oracle_mod.lib().fo_sum_potential_energies(e)
        if E0 is None:
            E0 = E
        err = max(err, abs(E - E0))
    return err / p.N


def test_leapfrog_energy_error_is_second_order(oracle_mod):
    e1 = _lf_energy_error(oracle_mod, 0.002, 0.2)
    e2 = _lf_energy_error(oracle_mod, 0.001, 0.2)
    assert e1 < 5e-3          # eV per unit cell, bounded over 100 steps
    assert 2.5 < e1 / e2 < 6.0


def _np_hamiltonian_error(oracle_mod, dt, t_total):
    p = batio3_zvr(6, kelvin=300.0)
    p.dt = dt
    o = oracle_mod.Oracle(p)
    _random_state(o, p, 31)
    e = o.md_prologue()
    assert e.H_Nose_Poincare == pytest.approx(0.0, abs=1e-12)
    hmax = 0.0
    for _ in range(int(round(t_total / dt))):
        o.nose_poincare(e)
        hmax = max(hmax, abs(e.H_Nose_Poincare) / p.N)
    return hmax, o.nose[0]


def test_nose_poincare_hamiltonian_stays_zero(oracle_mod):
    h1, s1 = _np_hamiltonian_error(oracle_mod, 0.002, 0.3)
    h2, s2 = _np_hamiltonian_error(oracle_mod, 0.001, 0.3)
    assert h1 < 1e-3                        # eV per unit cell from a far-from-equilibrium start
    assert 3.5 < h1 / h2 < 4.5              # second order in dt
    assert abs(s1 - 1.0) > 0.1              # the thermostat is active
    assert s1 == pytest.approx(s2, rel=1e-3)


def test_velocity_scaling_holds_target(oracle_mod):
    p = batio3_zvr(6, method="vs", kelvin=250.0)
    o = oracle_mod.Oracle(p)
    _random_state(o, p, 41)
    e = o.md_prologue()
    for _ in range(5):
        o.velocity_scaling(e)
        assert o.kinetic_energy_dipo() == pytest.approx(p.target_kinetic_energy, rel=1e-12)
