"""film / epit geometry on the device against the oracle.

mirror_force_mirror mirrors u into the upper half of z before the force
(mirror.F:9-36, mirror_force_mirror.F:20-27) and dV/du after it (:80-84);
read_defects_u.F:52-69 leaves the upper half uncounted and freezes the dead
layers; initialize_E_field.F:68-71 mirrors the external field, which the
per-step field updates of molecular-dynamics.F:66-76 then rewrite
unmirrored.  Tolerances as in test_gpu_parity.py: force and energies 1e-12,
trajectories 1e-10 normwise.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-10
ENERGY_TERMS = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain",
                "homo_coupling", "inho_strain", "inho_coupling")
E_STATIC = (0.0005, -0.0004, 0.002)
GAP_U = (0.0, 0.0, 0.02)


def _params(L, geometry, gap, method, acoustic=False, E=E_STATIC):
    p = batio3_zvr(L, method=method, kelvin=200.0, acoustic_mass_amu=46.44 if acoustic else None)
    p.bulk_or_film = geometry
    p.gap_id = gap
    p.gap_dipole_u = list(GAP_U)
    p.epi_strain = -0.005
    p.external_E_field = list(E)
    p.E_field_always_zero = False
    return p


def _dead(p):
    return [] if p.gap_id == 0 else ([0] if p.gap_id == 1 else [0, p.Lz // 2])


def _start(oracle_mod, p, seed):
    from feram_amd.system import System
    print(f"[gpu] {p.bulk_or_film} gap_id={p.gap_id} L=({p.Lx},{p.Ly},{p.Lz}) method={p.method}", flush=True)
    o = oracle_mod.Oracle(p)
    g = System(p)
    rng = np.random.default_rng(seed)
    u = 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    u[..., 2] += 0.08
    for z in _dead(p):
        u[z] = GAP_U                      # read_defects_u.F:56-69
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    if p.mass_acou[0] > 0.0:
        w = 0.01 * rng.standard_normal(u.shape)
        o.field("acouR")[...] = w
        g.put("acouR", w)
    return o, g


def _nw(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _check_energies(eg, eo, tol):
    scale = sum(abs(eo[t]) for t in ENERGY_TERMS)
    for t in ENERGY_TERMS:
        assert abs(eg[t] - eo[t]) <= tol * scale + tol * abs(eo[t]), (t, eg[t], eo[t])


CASES = [("epit", 0, "md", (8, 6, 16)), ("film", 0, "lf", (8, 8, 12)), ("epit", 2, "vs", (8, 6, 16)),
         ("film", 2, "md", (6, 8, 8)), ("epit", 0, "lf", (12, 8, 8)),
         # gap_id 1 needs an odd Lz (param_module.F:475-477): the 16x16x81 examples'
         ("epit", 1, "md", (16, 16, 81)), ("film", 1, "vs", (8, 8, 81)), ("epit", 1, "lf", (8, 12, 243))]


@pytest.mark.parametrize("geometry,gap,method,L", CASES)
def test_film_trajectory(oracle_mod, geometry, gap, method, L):
    p = _params(L, geometry, gap, method)
    o, g = _start(oracle_mod, p, 11 + gap)
    e_o = o.md_prologue()
    eg = g.md_prologue()
    assert _nw(g.get("dipoR"), o.field("dipoR")) == 0.0
    assert _nw(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    _check_energies(eg, e_o.as_dict(), FORCE_TOL)
    for chunk in (1, 5):
        for _ in range(chunk):
            o.step(e_o)
        eg = g.step(chunk)
        eo = e_o.as_dict()
        for name in ("dipoR", "dipoP", "dVddi"):
            assert _nw(g.get(name), o.field(name)) <= TRAJ_TOL, name
        _check_energies(eg, eo, TRAJ_TOL)
        assert eg["dipo_kinetic"] == pytest.approx(eo["dipo_kinetic"], rel=TRAJ_TOL)
        if method == "md":
            assert g.nose[0] == pytest.approx(o.nose[0], rel=TRAJ_TOL)
    u = g.get("dipoR")
    for z in _dead(p):
        assert np.array_equal(u[z], np.broadcast_to(GAP_U, u[z].shape))
    sign = np.array([-1.0, -1.0, 1.0])
    src = (p.Lz - 1) if gap == 0 else p.Lz
    lo = p.Lz - p.Lz // 2 if gap == 0 else p.Lz - (p.Lz + 1) // 2 + 1
    for z in range(lo, p.Lz):
        assert np.array_equal(u[z], sign * u[src - z])
    g.close()


def test_film_acoustic_leapfrog(oracle_mod):
    p = _params((8, 8, 12), "film", 0, "lf", acoustic=True)
    o, g = _start(oracle_mod, p, 5)
    e_o = o.md_prologue()
    g.md_prologue()
    for chunk in (1, 6):
        for _ in range(chunk):
            o.step(e_o)
        eg = g.step(chunk)
        for name in ("dipoR", "dipoP", "acouR", "acouP"):
            assert _nw(g.get(name), o.field(name)) <= TRAJ_TOL, name
        assert eg["acou_kinetic"] == pytest.approx(e_o.acou_kinetic, rel=TRAJ_TOL)
    g.close()


def test_film_field_unmirrored_after_per_step_update(oracle_mod):
    """A triangular wave: the prologue sees initialize_E_field's mirrored
    field, every step after it the unmirrored E_local + E_add."""
    p = _params((8, 6, 12), "epit", 0, "md", E=(0.0, 0.0, 0.0))
    o, g = _start(oracle_mod, p, 8)
    A, period, ntherm = [0.0005, 0.001, 0.003], 8, 2
    E0 = np.broadcast_to(oracle_mod.wave("triangular_cos", A, period, ntherm, 0, 0), (p.Lz, p.Ly, p.Lx, 3)).copy()
    E0[p.Lz // 2:] *= np.array([-1.0, -1.0, 1.0])
    o.field("E_field")[...] = E0
    g.set_E_schedule("triangular_cos", A, period, ntherm, 0, i_step=0)
    e_o = o.md_prologue()
    eg = g.md_prologue()
    assert eg["dipole_E_field"] == pytest.approx(e_o.dipole_E_field, rel=1e-11)
    g.set_E_schedule("triangular_cos", A, period, ntherm, 0, i_step=1)
    nsteps = 10
    for i in range(1, nsteps + 1):
        o.field("E_field")[...] = oracle_mod.wave("triangular_cos", A, period, ntherm, 0, i)
        o.nose_poincare(e_o)
    eg = g.nose_poincare(nsteps)
    assert _nw(g.get("dipoR"), o.field("dipoR")) <= TRAJ_TOL
    assert eg["dipole_E_field"] == pytest.approx(e_o.dipole_E_field, rel=1e-9, abs=1e-12)
    g.close()


@pytest.mark.parametrize("method", ["md", "vs"])
def test_film_batch_matches_individual_steps(method):
    """film, epit (gap_id 0 and 2) and bulk replicas of one size class in one
    batch evolve as their own step calls make them evolve."""
    from feram_amd.system import Batch, System
    print(f"[gpu] film batch {method}", flush=True)
    kinds = [("epit", 0), ("film", 2), ("bulk", 0), ("epit", 2)]

    def build():
        out = []
        for i, (geom, gap) in enumerate(kinds):
            p = _params((8, 8, 16), geom, gap, method)
            p.Q_Nose = 0.1 * p.N / 16 ** 3
            g = System(p)
            rng = np.random.default_rng(40 + i)
            u = 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
            for z in _dead(p):
                u[z] = GAP_U
            g.put("dipoR", u)
            g.md_prologue()
            out.append(g)
        return out

    single, batched = build(), build()
    es = [g.step(6) for g in single]
    with Batch(batched) as bt:
        assert bt.classes == 1
        eb = bt.step(6)
    for i, (a, b) in enumerate(zip(single, batched)):
        for f in ("dipoR", "dipoP", "dVddi"):
            assert _nw(b.get(f), a.get(f)) <= 1e-12, (i, f)
        scale = sum(abs(v) for v in es[i].values())
        for k, v in es[i].items():
            assert abs(eb[i][k] - v) <= 1e-12 * scale, (i, k)


FP32_TOL = 2e-5


@pytest.mark.parametrize("geometry,gap,method", [("epit", 0, "md"), ("film", 2, "lf")])
def test_film_fp32_trajectory(oracle_mod, geometry, gap, method):
    """FP32 mode (FP32 spectra and stored state): the film path against the
    FP64 oracle to the FP32 trajectory tolerance of the bulk tests."""
    from feram_amd.system import System
    p = _params((16, 16, 16), geometry, gap, method)
    print(f"[gpu] fp32 {geometry} gap_id={gap} {method}", flush=True)
    o = oracle_mod.Oracle(p)
    g = System(p, precision=1)
    rng = np.random.default_rng(23)
    u = rng.uniform(-0.05, 0.05, size=(p.Lz, p.Ly, p.Lx, 3))
    for z in _dead(p):
        u[z] = GAP_U
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    for _ in range(20):
        o.step(e_o)
    eg = g.step(20)
    for name in ("dipoR", "dipoP"):
        assert _nw(g.get(name), o.field(name)) <= FP32_TOL, name
    _check_energies(eg, e_o.as_dict(), FP32_TOL)
    g.close()


@pytest.mark.parametrize("method,f", [("md", 0.004), ("vs", -0.006), ("md", 0.0)])
def test_e110_trajectory(oracle_mod, method, f):
    """'e110' homogeneous strain (optimize-homo-strain.F:42-56, the reference's
    formula kept, test_oracle_e110 pins it): the device follows the oracle's
    trajectory, strain and energies."""
    from feram_amd.system import System
    p = batio3_zvr((12, 8, 16), method=method, kelvin=250.0)
    p.bulk_or_film = "e110"
    p.epi_strain = f
    p.GPa = [-4.0, -6.0, -2.0]
    print(f"[gpu] e110 {method} epi_strain={f}", flush=True)
    o = oracle_mod.Oracle(p)
    g = System(p)
    rng = np.random.default_rng(31)
    u = 0.06 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    eg = g.md_prologue()
    _check_energies(eg, e_o.as_dict(), FORCE_TOL)
    assert np.allclose(g.strain, o.strain, rtol=1e-12, atol=1e-16)
    for chunk in (1, 9):
        for _ in range(chunk):
            o.step(e_o)
        eg = g.step(chunk)
        for name in ("dipoR", "dipoP", "dVddi"):
            assert _nw(g.get(name), o.field(name)) <= TRAJ_TOL, name
        _check_energies(eg, e_o.as_dict(), TRAJ_TOL)
        assert np.allclose(g.strain, o.strain, rtol=TRAJ_TOL, atol=1e-14)
        assert g.strain[2] == f and g.strain[3] == 0.0 and g.strain[4] == 0.0
    g.close()
