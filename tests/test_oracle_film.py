"""Pins of the oracle's film/epit geometry (mirror.F:9-36,
mirror_force_mirror.F:20-27 and 80-84, read_defects_u.F:52-69,
initialize_E_field.F:68-71).

* the layer map of mirror(), written out by hand for the three gap_id cases;
* physics: a configuration whose upper half is the mirror image
  (u_x, u_y, u_z) -> (-u_x, -u_y, u_z) of its lower half is a fixed point of
  the z-reflection combined with u -> -u, a symmetry of the dipole-dipole,
  short-range, on-site and field terms (E mirrored with u), so without the
  strain coupling the *bulk* force on it already obeys the mirror relation,
  and a planted mirror with the wrong sign on u_z breaks it; with the
  coupling the film force equals the bulk force on the source layers and its
  mirror on the others;
* epit strain against the bulk 6x6 minimum;
* dead layers are frozen at gap_dipole_u and not counted; the upper half is
  not counted; the external field is mirrored.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr


def _film(L, geometry="epit", gap_id=0, method="lf", E=(0.0, 0.0, 0.0), gap_u=(0.0, 0.0, 0.0)):
    p = batio3_zvr(L, method=method, kelvin=150.0)
    p.bulk_or_film = geometry
    p.gap_id = gap_id
    p.gap_dipole_u = list(gap_u)
    p.epi_strain = -0.01
    if any(E):
        p.external_E_field = list(E)
        p.E_field_always_zero = False
    return p


# (Lz, gap_id) -> {target layer: source layer}, read off mirror.F:18-33 by hand
MAPS = {
    (4, 0): {3: 0, 2: 1},
    (6, 0): {5: 0, 4: 1, 3: 2},
    (5, 1): {4: 1, 3: 2},
    (7, 1): {6: 1, 5: 2, 4: 3},
    (6, 2): {5: 1, 4: 2},
    (8, 2): {7: 1, 6: 2, 5: 3},
}


@pytest.mark.parametrize("Lz,gap", sorted(MAPS))
def test_mirror_layer_map(oracle_mod, Lz, gap):
    p = _film((4, 4, Lz), gap_id=gap)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(Lz + 10 * gap)
    u = 0.05 * rng.standard_normal((Lz, 4, 4, 3))
    o.field("dipoR")[...] = u
    o.mirror_force_mirror()
    got = o.field("dipoR")
    m = MAPS[(Lz, gap)]
    for z in range(Lz):
        if z in m:
            want = u[m[z]] * np.array([-1.0, -1.0, 1.0])
        else:
            want = u[z]
        assert np.array_equal(got[z], want), z
    F = o.field("dVddi")
    for t, s in m.items():
        assert np.array_equal(F[t], F[s] * np.array([-1.0, -1.0, 1.0]))


@pytest.mark.parametrize("gap", [0, 1, 2])
def test_counted_and_dead_layers(oracle_mod, gap):
    Lz = {0: 8, 1: 9, 2: 8}[gap]
    p = _film((4, 4, Lz), gap_id=gap, gap_u=(0.01, -0.02, 0.03))
    o = oracle_mod.Oracle(p)
    counted = np.ones(Lz, bool)
    counted[(Lz + 1) // 2:] = False
    dead = [] if gap == 0 else ([0] if gap == 1 else [0, Lz // 2])
    counted[dead] = False
    assert np.array_equal(o.sum_p.any(axis=(1, 2)), counted)
    assert np.array_equal(o.sum_p.all(axis=(1, 2)), counted)
    minv = o.field("mass_dipo_inv")
    u = o.field("dipoR")
    for z in range(Lz):
        if z in dead:
            assert not minv[z].any() and np.allclose(u[z], [0.01, -0.02, 0.03])
        else:
            assert (minv[z] > 0).all()


def test_external_field_mirrored(oracle_mod):
    p = _film((4, 4, 8), E=(0.001, 0.002, -0.03))
    o = oracle_mod.Oracle(p)
    E = o.field("E_field")
    assert np.allclose(E[:4], [0.001, 0.002, -0.03])
    assert np.allclose(E[4:], [-0.001, -0.002, -0.03])


def _mirrored_config(L, seed, gap=0):
    Lx, Ly, Lz = L
    rng = np.random.default_rng(seed)
    u = 0.02 + 0.05 * rng.standard_normal((Lz, Ly, Lx, 3))
    sign = np.array([-1.0, -1.0, 1.0])
    if gap == 0:
        for z in range(Lz // 2):
            u[Lz - 1 - z] = sign * u[z]
    return u


def _bulk_force(p, u, E_field=None):
    import oracle
    o = oracle.Oracle(p)
    if E_field is not None:
        o.field("E_field")[...] = E_field
    o.field("dipoR")[...] = u
    o.mirror_force_mirror()
    return o.field("dVddi").copy()


def test_bulk_force_obeys_the_mirror_relation_without_strain_coupling():
    """Dipole-dipole, short-range, on-site and field terms are symmetric
    under the mirror; the inhomogeneous strain stencil (forward differences
    of w) is not, which is why mirror_force_mirror mirrors dV/du explicitly."""
    L = (8, 6, 8)
    p = _film(L, E=(0.0005, -0.0004, 0.002))
    p.bulk_or_film = "bulk"
    p.B1xx = p.B1yy = p.B4yz = 0.0
    E = np.broadcast_to(np.array([0.0005, -0.0004, 0.002]), (L[2], L[1], L[0], 3)).copy()
    E[4:] *= np.array([-1.0, -1.0, 1.0])
    Fb = _bulk_force(p, _mirrored_config(L, 3), E)
    sign = np.array([-1.0, -1.0, 1.0])
    dev = max(np.max(np.abs(Fb[L[2] - 1 - z] - sign * Fb[z])) for z in range(L[2] // 2))
    assert dev <= 1e-12 * np.max(np.abs(Fb))


def test_film_force_is_mirrored_bulk_force(oracle_mod):
    """film (bulk homogeneous strain): dV/du on the source layers is the bulk
    force of the mirrored configuration, the mirrored layers carry its mirror."""
    L = (8, 6, 8)
    E = (0.0005, -0.0004, 0.002)
    pf = _film(L, geometry="film", E=E)
    of = oracle_mod.Oracle(pf)
    u = _mirrored_config(L, 3)
    of.field("dipoR")[...] = u
    Ef = of.field("E_field").copy()
    of.mirror_force_mirror()
    Ff = of.field("dVddi")
    pb = _film(L, geometry="film", E=E)
    pb.bulk_or_film = "bulk"
    Fb = _bulk_force(pb, u, Ef)
    sign = np.array([-1.0, -1.0, 1.0])
    assert np.array_equal(Ff[:4], Fb[:4])
    for z in range(4):
        assert np.array_equal(Ff[7 - z], sign * Fb[z])


def test_planted_wrong_mirror_sign_breaks_the_symmetry(oracle_mod):
    """A mirror that flipped u_z instead of keeping it leaves the bulk force
    without the mirror relation the test above relies on."""
    L = (8, 6, 8)
    pb = _film(L)
    pb.bulk_or_film = "bulk"
    pb.B1xx = pb.B1yy = pb.B4yz = 0.0
    u = _mirrored_config(L, 3)
    for z in range(L[2] // 2):
        u[L[2] - 1 - z] = u[z] * np.array([-1.0, -1.0, -1.0])      # planted
    Fb = _bulk_force(pb, u)
    sign = np.array([-1.0, -1.0, 1.0])
    dev = max(np.max(np.abs(Fb[L[2] - 1 - z] - sign * Fb[z])) for z in range(L[2] // 2))
    assert dev > 1e-3 * np.max(np.abs(Fb))


def test_epit_strain_agrees_with_bulk_minimum(oracle_mod):
    """epit (optimize-homo-strain.F:34-41) clamps eta_1 = eta_2 = epi_strain,
    eta_6 = 0 and minimises over eta_3..5.  With u along z only, the bulk
    minimum over all six (a separate 6x6 solve) has eta_1 = eta_2 and
    eta_6 = 0; clamping epi_strain at that eta_1 must give the bulk eta_3..5."""
    L = (6, 6, 8)
    u = _mirrored_config(L, 5)
    u[..., :2] = 0.0
    pb = _film(L)
    pb.bulk_or_film = "bulk"
    ob = oracle_mod.Oracle(pb)
    ob.field("dipoR")[...] = u
    ob.mirror_force_mirror()
    sb = ob.strain.copy()
    assert sb[0] == pytest.approx(sb[1], rel=1e-12) and abs(sb[5]) < 1e-15
    pe = _film(L)
    pe.epi_strain = float(sb[0])
    oe = oracle_mod.Oracle(pe)
    oe.field("dipoR")[...] = u
    oe.mirror_force_mirror()
    se = oe.strain.copy()
    assert se[0] == se[1] == pe.epi_strain and se[5] == 0.0
    assert np.allclose(se[2:5], sb[2:5], rtol=1e-10, atol=1e-15)
    # a different clamp moves eta_3 (Poisson effect)
    pe.epi_strain = float(sb[0]) - 0.01
    oe2 = oracle_mod.Oracle(pe)
    oe2.field("dipoR")[...] = u
    oe2.mirror_force_mirror()
    assert oe2.strain[2] > se[2] + 1e-4
