"""The oracle's add_to_Average (average_module.F:57-163) against closed forms.

Uniform dipoles make every sum a product (U = n u, P = Z* n u, the products of
the sums, one histogram bin per component holding every counted cell); the
sum_p mask must drop the uncounted cells from every sum and histogram; the
histogram bins follow Fortran NINT (half away from zero) at step 0.005; the
energy sums, total**2 and KE / s**2 follow from the energies passed in.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr


def _oracle(oracle_mod, L=4):
    p = batio3_zvr(L)
    return oracle_mod.Oracle(p, tables=False), p


def _energy(oracle_mod, total, kin):
    e = oracle_mod.FoEnergy()
    e.total = total
    e.dipo_kinetic = kin
    e.short_range = 0.25
    e.H_Nose_Poincare = -1.5
    return e


def test_uniform_dipoles(oracle_mod):
    o, p = _oracle(oracle_mod)
    u = np.array([0.031, -0.0125, 0.0024])
    o.field("dipoR")[...] = u
    o.field("Z_u_R")[...] = p.Z_star[0] * u
    o.nose[0] = 1.25
    o.strain[:] = [1, 2, 3, 4, 5, 6]
    acc = oracle_mod.FoAverage()
    d1 = np.zeros((199, 3), np.int64)
    d3 = np.zeros((199, 199, 199), np.int32)
    for _ in range(3):
        o.add_to_average(_energy(oracle_mod, 2.0, 0.5), acc, d1, d3)
    a = acc.as_dict()
    n = p.N
    U = n * u
    assert a["n_add"] == 3
    assert a["u"] == pytest.approx(3 * U, rel=1e-14)
    assert a["uu"] == pytest.approx(3 * np.array([U[0] ** 2, U[1] ** 2, U[2] ** 2, U[1] * U[2], U[2] * U[0], U[0] * U[1]]),
                                    rel=1e-14)
    P = p.Z_star[0] * U
    assert a["p"] == pytest.approx(3 * P, rel=1e-14)
    assert a["pp"][3] == pytest.approx(3 * P[1] * P[2], rel=1e-14)
    assert a["uu_tmp"] == pytest.approx(n * u ** 2, rel=1e-14)
    assert a["e2"] == pytest.approx(3 * 4.0)
    assert a["dipo_kinetic_true"] == pytest.approx(3 * 0.5 / 1.25 ** 2)
    assert a["strain"] == pytest.approx([3, 6, 9, 12, 15, 18])
    assert a["energy"]["short_range"] == pytest.approx(0.75)
    assert a["energy"]["H_Nose_Poincare"] == pytest.approx(-4.5)
    # nint(0.031*200) = 6, nint(-0.0125*200) = nint(-2.5) = -3, nint(0.0024*200) = nint(0.48) = 0
    for q, k in enumerate((6, -3, 0)):
        assert d1[k + 99, q] == 3 * n
        assert d1[:, q].sum() == 3 * n
    assert d3[0 + 99, -3 + 99, 6 + 99] == 3 * n     # [kz][ky][kx]
    assert d3.sum() == 3 * n


def test_sum_p_mask_and_range(oracle_mod):
    o, p = _oracle(oracle_mod)
    rng = np.random.default_rng(5)
    u = rng.uniform(-0.2, 0.2, size=(p.Lz, p.Ly, p.Lx, 3))
    u[0, 0, 0] = [0.6, 0.0, 0.0]          # outside -99..99 bins along x: not in dist3, x not in dist1
    o.field("dipoR")[...] = u
    o.field("Z_u_R")[...] = p.Z_star[0] * u
    mask = rng.integers(0, 2, size=(p.Lz, p.Ly, p.Lx)).astype(np.uint8)
    mask[0, 0, 0] = 1
    o.sum_p[...] = mask
    acc = oracle_mod.FoAverage()
    d1 = np.zeros((199, 3), np.int64)
    d3 = np.zeros((199, 199, 199), np.int32)
    o.add_to_average(_energy(oracle_mod, 1.0, 0.0), acc, d1, d3)
    a = acc.as_dict()
    cnt = mask.astype(bool)
    assert a["u_tmp"] == pytest.approx(u[cnt].sum(axis=0), rel=1e-13)
    assert a["uu_tmp"] == pytest.approx((u[cnt] ** 2).sum(axis=0), rel=1e-13)
    k = np.rint(u[cnt] * 200)            # no value sits on a half step here
    inr = np.abs(k) <= 99
    for q in range(3):
        assert d1[:, q].sum() == inr[:, q].sum()
    assert d3.sum() == inr.all(axis=1).sum()
