"""The oracle against the known answers the reference's own tests assert.

Pins: modulation_check.sh (elastic tables, c_k_modulation, optimize_inho_strain,
the inhomogeneous energies and the c2r of the acoustic field),
dispersion_check.sh (Ewald long-range kernel at Gamma to 15 digits; the
inhomogeneous elastic matrix K(k)), marsaglia_tsang_uni64_check.F (the seeded
RNG behind initialize_dipoR's normal initialisation).
"""
import numpy as np
import pytest

from conftest import param_from_dict


def f_e(v, mant, expw):
    """Fortran Ew.d: 0.ddd...E+xx with `mant` mantissa digits and `expw` exponent digits."""
    if v == 0.0:
        return " 0." + "0" * mant + "E+" + "0" * expw
    s = "%.*E" % (mant - 1, abs(v))
    m, ex = s.split("E")
    ex = int(ex) + 1
    return ("-" if v < 0 else " ") + "0." + m.replace(".", "") + "E%+0*d" % (expw + 1, ex)


def test_modulation_check(goldens, oracle_mod):
    g = goldens["modulation_check"]
    p = param_from_dict(g["params"])
    mod = np.zeros((p.Lz, p.Ly, p.Lx))
    for x, y, z, a in g["modulation"]:
        mod[z, y, x] = a
    o = oracle_mod.Oracle(p, modulation_R=mod)
    e = o.md_prologue()
    ex = g["expect"]
    # the log prints step 0, 1 and 2 (n_thermalize = n_average = 1, verbose = 2)
    lines_is, lines_im = 0, 0
    for step in range(3):
        if step > 0:
            o.leapfrog(e, step)
        lines_is += "%.12f" % (e.inho_strain / p.N) == ex["inho_strain_per_uc_f18_12"]
        lines_im += "%.12f" % (e.inho_modulation / p.N) == ex["inho_modulation_per_uc_f18_12"]
        if step == 1:
            # write_system at step 1 reconstructs acouR = c2r(-i acouK)/N (coord_module.F:457-469)
            aK = o.field("acouK")
            acouR = np.stack([o.c2r(-1j * aK[..., a]) for a in range(3)], -1) / p.N
            cells = acouR.reshape(-1, 3)
            pos = sum(any(f_e(v, 6, 2).rjust(14) == ex["acouR_e14_6_positive"] for v in c) for c in cells)
            neg = sum(any(f_e(v, 6, 2).rjust(14) == ex["acouR_e14_6_negative"] for v in c) for c in cells)
            assert pos == ex["acouR_positive_lines"]
            assert neg == ex["acouR_negative_lines"]
    assert lines_is == ex["inho_strain_log_lines"]
    assert lines_im == ex["inho_modulation_log_lines"]


def _print_path(L):
    """print_eigenvalues' k path for Lx=Ly=Lz (print-eigenvalues.F:19-50)."""
    h = L // 2
    pts = [(ix, 0, 0) for ix in range(0, h + 1)]
    pts += [(h, iy, 0) for iy in range(1, h + 1)]
    pts += [(ix, ix, 0) for ix in range(h - 1, -1, -1)]
    pts += [(ix, ix, ix) for ix in range(1, h + 1)]
    pts += [(h, iy, iy) for iy in range(h - 1, -1, -1)]
    pts += [(h, iy, iy) for iy in range(1, h + 1)]
    pts += [(h, h, iz) for iz in range(h - 1, -1, -1)]
    return pts


def test_dispersion_check(goldens, oracle_mod):
    g = goldens["dispersion_check"]
    ex = g["expect"]
    n_long, n_K = 0, 0
    for L in g["L_values"]:
        p = param_from_dict(g["params"], Lx=L, Ly=L, Lz=L)
        o = oracle_mod.Oracle(p)
        lr = o.field("long_range")
        K = o.field("inhomo_K")
        Z2 = p.Z_star[0] ** 2
        for (ix, iy, iz) in _print_path(L):
            ev = np.linalg.eigvalsh(Z2 * lr[iz, iy, ix].T)
            n_long += ex["long_dat_digits"] in " ".join(f_e(v, 16, 3) for v in ev)
            ev = np.linalg.eigvalsh(K[iz, iy, ix].T)
            n_K += ex["inhomo_K_dat_digits"] in " ".join(f_e(v, 16, 3) for v in ev)
    assert n_long == ex["long_dat_lines"]
    assert n_K == ex["inhomo_K_dat_lines"]


@pytest.mark.slow
def test_marsaglia_tsang_known_answers(goldens, oracle_mod):
    g = goldens["marsaglia_tsang_uni64_check"]
    o = oracle_mod
    first = o.uni64_seed(*g["seed"])
    assert first >= 1.0 and float(first).is_integer()
    o.uni64_fill(g["skip"])
    got = o.uni64_fill(5) * 9007199254740992.0
    assert list(got) == g["next_five_times_2pow53"]
    # re-seeding returns the running count of initialisations
    n = o.normal_dist_seed(*g["normal_seed_pairs"][0])
    assert n == first + 1.0
    vals = np.array([o.normal_dist(50.0, 10.0) for _ in range(20000)])
    assert abs(vals.mean() - 50.0) < 0.3 and abs(vals.std() - 10.0) < 0.3
