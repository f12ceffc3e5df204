"""The oracle's FFT against brute-force DFTs and an independent library.

The reference's own FFT test only pins a DC round trip
(feram_fftw_wisdom_in.F:57-89); these checks pin general spectra, including
the c2r rule for non-Hermitian half spectra (Im of kx=0 and kx=Lx/2 rows is
ignored after the y/z inverse passes), which the inhomogeneous-strain path
exercises on its Nyquist planes.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr


def dft(x, sign):
    n = len(x)
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 14, 16, 24, 27, 32, 48, 96, 128])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_matches_brute_force(oracle_mod, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = oracle_mod.fft1d(x, sign)
    ref = dft(x, sign)
    assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("L", [(4, 4, 4), (6, 4, 8), (8, 12, 6), (16, 16, 16)])
def test_r2c_matches_numpy(oracle_mod, L):
    p = batio3_zvr(L)
    o = oracle_mod.Oracle(p, tables=False)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((p.Lz, p.Ly, p.Lx))
    got = o.r2c(f)
    ref = np.fft.rfftn(f, axes=(0, 1, 2))     # forward e^{-i}, x halved
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    back = o.c2r(got) / p.N
    assert np.max(np.abs(back - f)) <= 1e-13 * np.max(np.abs(f))


def c2r_rule(spec, Lx):
    """out(R) = sum_kx c(kx) Re[e^{i kx x} sum_{ky,kz} X e^{i(ky y + kz z)}]."""
    Lz, Ly, Kx = spec.shape
    yz = np.fft.ifftn(spec, axes=(0, 1)) * (Ly * Lz)     # exact complex inverse along z, y
    x = np.arange(Lx)
    out = np.zeros((Lz, Ly, Lx))
    for kx in range(Kx):
        c = 1.0 if kx == 0 or (Lx % 2 == 0 and kx == Lx // 2) else 2.0
        out += c * np.real(yz[:, :, kx, None] * np.exp(2j * np.pi * kx * x / Lx)[None, None, :])
    return out


@pytest.mark.parametrize("L", [(4, 6, 8), (8, 8, 8)])
def test_c2r_non_hermitian_rule(oracle_mod, L):
    p = batio3_zvr(L)
    o = oracle_mod.Oracle(p, tables=False)
    rng = np.random.default_rng(2)
    spec = rng.standard_normal((p.Lz, p.Ly, p.Lx // 2 + 1)) + 1j * rng.standard_normal(
        (p.Lz, p.Ly, p.Lx // 2 + 1))
    got = o.c2r(spec)
    ref = c2r_rule(spec, p.Lx)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
