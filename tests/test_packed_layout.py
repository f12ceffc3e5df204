"""The packed half-spectrum layout and its k-space evaluation plan, modelled in numpy.

The device keeps the real kx = 0 and kx = Lx/2 values of every x row in one
complex column, splits that column into P / Q pencils in the y pass and into
P / Q points in the z pass, and evaluates the k-space stage once per stored
point, or at the point and its mirror partner where the partner's k vector is
not exactly -k.  This model follows the same steps (same split/join formulas,
same plan as kplan in fhip_kernels.hpp) and must reproduce what the reference
computes on its full half spectrum: rfftn, a per-point kernel, and a c2r that
keeps the Hermitian part (FFTW's rule, reproduced by numpy's irfft).  The
kernels are chosen so that a dropped partner evaluation, a wrong partner index
or a swapped P/Q slot changes the result: the strain kernel is a function of
the folded k vector that is not Hermitian where the partner is not exactly
-k, and the dipole-dipole kernel is a table over the stored points that is
not symmetric under k -> -k (as the reference's L(k) is not exactly, on
anisotropic supercells, from the asymmetric image ranges of its Ewald sum).
"""
import numpy as np
import pytest


def kcomp(i, L, fold):
    f = i / L
    if fold and f > 0.5:
        f -= 1.0
    return f * 2 * np.pi


def kvec(kx, ky, kz, Lx, Ly, Lz):
    return np.array([kcomp(kx, Lx, False), kcomp(ky, Ly, True), kcomp(kz, Lz, True)])


def strain_kernel(shape):
    """Kernel of the folded k vector with an odd imaginary part: Hermitian-
    preserving only where the partner is exactly -k."""
    Lz, Ly, Lx = shape

    def g(p):
        k = kvec(*p, Lx, Ly, Lz)
        return 1 + k[0] ** 2 + 0.3 * k[1] * k[2] + 1j * (k[0] + 0.5 * k[1] - k[2] + 0.2 * k[0] * k[1] * k[2])
    return g


def dd_kernel(shape, seed=11):
    """A real table over the half-spectrum points, not symmetric under k -> -k."""
    Lz, Ly, Lx = shape
    table = 1 + np.random.default_rng(seed).uniform(0.0, 1.0, size=(Lx // 2 + 1, Ly, Lz))
    return lambda p: table[p]


def reference(a, g):
    Lz, Ly, Lx = a.shape
    H = np.fft.fft(np.fft.fft(np.fft.rfft(a, axis=2), axis=1), axis=0)
    O = np.empty_like(H)
    E = 0.0
    for kz in range(Lz):
        for ky in range(Ly):
            for kx in range(Lx // 2 + 1):
                gk = g((kx, ky, kz))
                O[kz, ky, kx] = gk * H[kz, ky, kx]
                wgt = 1.0 if kx in (0, Lx // 2) else 2.0
                E += wgt * np.real(np.conj(H[kz, ky, kx]) * gk * H[kz, ky, kx])
    out = np.fft.irfft(np.fft.ifft(np.fft.ifft(O, axis=0), axis=1), n=Lx, axis=2)
    return out, E


def split_pq(v, m, lower):
    return (v + np.conj(m)) / 2 if lower else (m - np.conj(v)) / 2j


def join_pq(v, m, lower):
    return v + 1j * m if lower else np.conj(m) + 1j * np.conj(v)


def kplan(s, t, Lx, Ly, Lz, dd):
    """Evaluations of the element (kx = 0 pencil, slot s, index t): list of (point, mode), weight.
    The dipole-dipole group evaluates both points of every pair; the strain
    group only where the partner's folded k vector is not exactly -k."""
    Mx = Lx // 2
    tp = 0 if t == 0 else Lz - t
    if s != 0 and 2 * s < Ly:
        if dd or 2 * t == Lz:
            return [((0, s, t), 'N'), ((0, Ly - s, tp), 'C')], 0.5
        return [((0, s, t), 'N')], 1.0
    if 2 * s > Ly:
        return [((Mx, Ly - s, t), 'N'), ((Mx, s, tp), 'C')], 0.5
    if t == 0 or 2 * t == Lz:
        return [((0, s, t), 'RP'), ((Mx, s, t), 'RQ')], 0.5
    if 2 * t < Lz:
        if dd or s != 0:
            return [((0, s, t), 'N'), ((0, s, tp), 'C')], 0.5
        return [((0, s, t), 'N')], 1.0
    return [((Mx, s, Lz - t), 'N'), ((Mx, s, t), 'C')], 0.5


def packed_pipeline(a, g, dd, plan_fn=kplan):
    Lz, Ly, Lx = a.shape
    Mx = Lx // 2
    X = np.fft.fft(a, axis=2)
    S = X[:, :, :Mx].copy()
    S[:, :, 0] = X[:, :, 0].real + 1j * X[:, :, Mx].real          # A[0] + i A[Lx/2]
    S = np.fft.fft(S, axis=1)
    C = S[:, :, 0].copy()
    for t in range(1, Ly):
        if 2 * t != Ly:
            S[:, t, 0] = split_pq(C[:, t], C[:, Ly - t], 2 * t < Ly)
    S = np.fft.fft(S, axis=0)
    rpen = (0, Ly // 2) if Ly % 2 == 0 else (0,)                    # R pencils
    for s in rpen:
        R = S[:, s, 0].copy()
        for t in range(1, Lz):
            if 2 * t != Lz:
                S[t, s, 0] = split_pq(R[t], R[Lz - t], 2 * t < Lz)
    Out = np.zeros_like(S)
    E = 0.0
    for t in range(Lz):
        for s in range(Ly):
            for kx in range(Mx):
                x = S[t, s, kx]
                if kx != 0:
                    gk = g((kx, s, t))
                    Out[t, s, kx] = gk * x
                    E += 2.0 * np.real(np.conj(x) * gk * x)
                    continue
                plan, wm = plan_fn(s, t, Lx, Ly, Lz, dd)
                half = 0.5 if len(plan) == 2 else 1.0
                out = 0j
                for (p, md) in plan:
                    xe = {'N': x, 'C': np.conj(x), 'RP': x.real + 0j, 'RQ': x.imag + 0j}[md]
                    o = g(p) * xe
                    E += 2.0 * wm * np.real(np.conj(xe) * o)
                    if md == 'N':
                        out += half * o
                    elif md == 'C':
                        out += half * np.conj(o)
                    elif md == 'RP':
                        out = o.real + 1j * out.imag
                    else:
                        out = out.real + 1j * o.real
                Out[t, s, kx] = out
    for s in rpen:
        R = Out[:, s, 0].copy()
        for t in range(1, Lz):
            if 2 * t != Lz:
                Out[t, s, 0] = join_pq(R[t], R[Lz - t], 2 * t < Lz)
    Out = np.fft.ifft(Out, axis=0)
    C = Out[:, :, 0].copy()
    for t in range(1, Ly):
        if 2 * t != Ly:
            Out[:, t, 0] = join_pq(C[:, t], C[:, Ly - t], 2 * t < Ly)
    Out = np.fft.ifft(Out, axis=1)
    Z = np.zeros((Lz, Ly, Mx + 1), complex)
    Z[:, :, 1:Mx] = Out[:, :, 1:Mx]
    Z[:, :, 0] = Out[:, :, 0].real
    Z[:, :, Mx] = Out[:, :, 0].imag
    return np.fft.irfft(Z, n=Lx, axis=2), E


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 12, 8), (4, 4, 4), (8, 6, 12), (12, 8, 16), (9, 6, 8), (8, 9, 6)])
@pytest.mark.parametrize("dd", [False, True])
def test_packed_pipeline_matches_half_spectrum(shape, dd):
    a = np.random.default_rng(7).standard_normal(shape)
    g = dd_kernel(shape) if dd else strain_kernel(shape)
    ref, Eref = reference(a, g)
    got, E = packed_pipeline(a, g, dd)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert abs(E - Eref) <= 1e-13 * abs(Eref)


def _first_only(s, t, Lx, Ly, Lz, dd):
    plan, _ = kplan(s, t, Lx, Ly, Lz, dd)
    return plan[:1], 1.0


def test_single_evaluation_of_the_nyquist_plane_is_wrong():
    """Evaluating the kx = Lx/2 points once (no partner) must fail: the check has teeth."""
    a = np.random.default_rng(3).standard_normal((8, 8, 8))
    g = strain_kernel(a.shape)
    ref, _ = reference(a, g)
    got, _ = packed_pipeline(a, g, False, plan_fn=_first_only)
    assert np.max(np.abs(got - ref)) > 1e-6 * np.max(np.abs(ref))


def test_single_dipole_evaluation_misses_the_table_asymmetry():
    """One evaluation per kx = 0 / Lx/2 pair with the point's own table value
    (a plan that assumes L(-k) = L(k)) misses the reference's result."""
    a = np.random.default_rng(5).standard_normal((8, 6, 12))
    g = dd_kernel(a.shape)
    ref, _ = reference(a, g)
    got, _ = packed_pipeline(a, g, True, plan_fn=_first_only)
    assert np.max(np.abs(got - ref)) > 1e-6 * np.max(np.abs(ref))
