"""HIP path against the FP64 oracle on identical inputs (needs an MI355X).

Tolerances, FP64 plain mode:
  * forces dV/du and dV/dw: normwise  max|dF| / max|F|          <= 1e-12
  * each energy term:       |dE| <= 1e-12 * sum of |terms|  (+ 1e-12 relative)
  * tables L(k):            max|dL| / max|L|                   <= 1e-12
  * trajectories (20 steps): u, p normwise                     <= 1e-10
The FFTs, the order of the Parseval and real-space sums and FMA contraction
differ between the two sides, so agreement is to rounding, not bitwise.
"""
import numpy as np
import pytest

from conftest import load_golden, param_from_dict
from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
ENERGY_TOL = 1e-12
TRAJ_TOL = 1e-10

ENERGY_TERMS = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain",
                "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation")


def _system(p):
    from feram_amd.system import System
    print(f"[gpu] System L=({p.Lx},{p.Ly},{p.Lz}) method={p.method}", flush=True)
    return System(p)


def _pair(oracle_mod, p, modulation=None):
    o = oracle_mod.Oracle(p, modulation_R=modulation)
    g = _system(p)
    if modulation is not None:
        g.set_modulation(modulation)
    return o, g


def _random(p, seed, amp=0.06, mean=0.08):
    rng = np.random.default_rng(seed)
    return mean + amp * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))


def _normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _check_energies(eg, eo, tol=ENERGY_TOL):
    scale = sum(abs(eo[t]) for t in ENERGY_TERMS)
    for t in ENERGY_TERMS:
        assert abs(eg[t] - eo[t]) <= tol * scale + tol * abs(eo[t]), (t, eg[t], eo[t])


@pytest.mark.parametrize("L", [8, 12, 16, (8, 12, 16), (16, 6, 24)])
def test_long_range_table(oracle_mod, L):
    p = batio3_zvr(L)
    o, g = _pair(oracle_mod, p)
    lo = o.field("long_range")                          # (z, y, kx, b, a)
    lg = g.get_long_range()[:, :p.Ly]
    assert _normwise(lg, lo) <= FORCE_TOL


@pytest.mark.parametrize("L", [8, 12, 16, (8, 12, 16), (16, 6, 24), 24, (8, 8, 256), (12, 6, 128), (6, 16, 64),
                               (16, 16, 80), (32, 20, 60), (40, 90, 12), (100, 16, 8), (16, 16, 81), (8, 20, 243),
                               (6, 81, 64)])
def test_force_and_energies(oracle_mod, L):
    p = batio3_zvr(L)
    o, g = _pair(oracle_mod, p)
    u = _random(p, 7)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    _check_energies(eg, eo)
    assert np.allclose(g.strain, o.strain, rtol=1e-12, atol=1e-16)


def test_force_with_fields_cells_and_pressure(oracle_mod):
    p = batio3_zvr((12, 8, 16))
    p.GPa = [-3.0, 1.0, 2.0]
    p.external_E_field = [0.001, -0.002, 0.0005]
    p.E_field_always_zero = False
    p.P_kappa2 = [5.502, 4.9, 6.1]
    p.P_alpha = [110.4, 100.0, 120.0]
    p.P_k1 = [0.0, 3.0, -2.0]
    p.P_k2 = [0.0, 1.0, 2.0]
    p.P_k3 = [0.0, -1.0, 4.0]
    p.P_k4 = [0.0, 5.0, 7.0]
    p.mass_dipo = [p.mass_dipo[0], 1.2 * p.mass_dipo[0], 0.8 * p.mass_dipo[0]]
    o, g = _pair(oracle_mod, p)
    rng = np.random.default_rng(3)
    relax = rng.integers(1, 4, size=(p.Lz, p.Ly, p.Lx)).astype(np.int32)
    o.relaxor[...] = relax
    for s in range(3):
        o.field("mass_dipo_inv")[relax == s + 1] = 1.0 / p.mass_dipo[s]
    g.set_cells(relaxor=relax)
    eloc = 0.002 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("E_field")[...] = eloc + np.array(p.external_E_field)
    g.put("E_field_local", eloc)
    u = _random(p, 8)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    _check_energies(eg, eo)


def test_acoustic_force(oracle_mod):
    p = batio3_zvr((8, 12, 8), method="lf", acoustic_mass_amu=46.44)
    o, g = _pair(oracle_mod, p)
    u = _random(p, 9)
    w = 0.02 * np.random.default_rng(10).standard_normal(u.shape)
    for name, v in (("dipoR", u), ("acouR", w)):
        o.field(name)[...] = v
        g.put(name, v)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    assert _normwise(g.get("dVdacR"), o.field("dVdacR")) <= FORCE_TOL
    _check_energies(eg, eo)


@pytest.mark.parametrize("method,L", [("md", 12), ("lf", 12), ("vs", 12), ("md", 40), ("vs", (20, 16, 81))])
def test_trajectory(oracle_mod, method, L):
    p = batio3_zvr(L, method=method, kelvin=300.0)
    o, g = _pair(oracle_mod, p)
    u = _random(p, 21)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    eg = g.md_prologue()
    _check_energies(eg, e_o.as_dict())
    if method == "md":
        assert g.nose == pytest.approx(tuple(o.nose), rel=1e-12, abs=1e-12)
    for chunk in (1, 4, 15):
        for i in range(chunk):
            o.step(e_o)
        eg = g.step(chunk)
        eo = e_o.as_dict()
        assert _normwise(g.get("dipoR"), o.field("dipoR")) <= TRAJ_TOL
        assert _normwise(g.get("dipoP"), o.field("dipoP")) <= TRAJ_TOL
        assert _normwise(g.get("dVddi"), o.field("dVddi")) <= TRAJ_TOL
        _check_energies(eg, eo, tol=TRAJ_TOL)
        assert eg["dipo_kinetic"] == pytest.approx(eo["dipo_kinetic"], rel=TRAJ_TOL)
        if method == "md":
            s_g, pi_g, _ = g.nose
            assert s_g == pytest.approx(o.nose[0], rel=TRAJ_TOL)
            assert pi_g == pytest.approx(o.nose[1], rel=1e-8, abs=1e-10)
            assert eg["H_Nose_Poincare"] == pytest.approx(eo["H_Nose_Poincare"], abs=1e-8 * p.N)


@pytest.mark.parametrize("method", ["lf", "vs"])
def test_acoustic_trajectory(oracle_mod, method):
    p = batio3_zvr((8, 8, 12), method=method, kelvin=250.0, acoustic_mass_amu=46.44)
    o, g = _pair(oracle_mod, p)
    u = _random(p, 31)
    w = 0.02 * np.random.default_rng(32).standard_normal(u.shape)
    for name, v in (("dipoR", u), ("acouR", w)):
        o.field(name)[...] = v
        g.put(name, v)
    e_o = o.md_prologue()
    g.md_prologue()
    for chunk in (1, 6):
        for _ in range(chunk):
            o.step(e_o)
        eg = g.step(chunk)
        for name in ("dipoR", "dipoP", "acouR", "acouP"):
            assert _normwise(g.get(name), o.field(name)) <= TRAJ_TOL, name
        assert eg["acou_kinetic"] == pytest.approx(e_o.acou_kinetic, rel=TRAJ_TOL)


def test_modulation_golden():
    g_ = load_golden()["modulation_check"]
    p = param_from_dict(g_["params"])
    mod = np.zeros((p.Lz, p.Ly, p.Lx))
    for x, y, z, a in g_["modulation"]:
        mod[z, y, x] = a
    g = _system(p)
    g.set_modulation(mod)
    ex = g_["expect"]
    e = g.md_prologue()
    assert "%.12f" % (e["inho_strain"] / p.N) == ex["inho_strain_per_uc_f18_12"]
    assert "%.12f" % (e["inho_modulation"] / p.N) == ex["inho_modulation_per_uc_f18_12"]
    e = g.leapfrog(1, 1)
    w = g.get("acouR_optimised").reshape(-1, 3)
    from test_oracle_goldens import f_e
    pos = sum(any(f_e(v, 6, 2).rjust(14) == ex["acouR_e14_6_positive"] for v in c) for c in w)
    neg = sum(any(f_e(v, 6, 2).rjust(14) == ex["acouR_e14_6_negative"] for v in c) for c in w)
    assert (pos, neg) == (ex["acouR_positive_lines"], ex["acouR_negative_lines"])
    assert "%.12f" % (e["inho_strain"] / p.N) == ex["inho_strain_per_uc_f18_12"]


def test_padded_layout_roundtrip():
    for pad in (1, 3):
        p = batio3_zvr((8, 12, 6))
        p.padding_y = pad
        g = _system(p)
        shp = g.padded_shape()
        host = np.random.default_rng(pad).standard_normal(shp)
        g.put_padded("dipoR", host)
        back = np.full(shp, 7.0)
        g.get_padded("dipoR", back)
        real = (slice(None), slice(0, p.Ly), slice(0, p.Lx))
        assert np.array_equal(back[real], host[real])
        mask = np.ones(shp, bool)
        mask[real] = False
        assert np.all(back[mask] == 7.0)          # padding untouched
        u = g.get("dipoR")
        assert np.array_equal(u, host[real])


def test_bitwise_repeatable():
    p = batio3_zvr(16)
    runs = []
    for _ in range(2):
        g = _system(p)
        g.put("dipoR", _random(p, 41))
        g.md_prologue()
        e = g.nose_poincare(5)
        runs.append((g.get("dipoR"), g.get("dipoP"), e["total"], g.nose))
        g.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2] and runs[0][3] == runs[1][3]


def test_rejects_unsupported():
    from feram_amd.system import FhipError
    p = batio3_zvr((10, 8, 8))
    with pytest.raises(FhipError, match="supercell edges"):
        _system(p)
    p = batio3_zvr(8)
    p.bulk_or_film = "film"
    p.gap_id = 1                         # needs odd Lz (param_module.F:475-477)
    with pytest.raises(FhipError, match="gap_id"):
        _system(p)
    p = batio3_zvr(8)
    p.gap_id = 2                         # dead layers only in film/epit
    with pytest.raises(FhipError, match="gap_id"):
        _system(p)


@pytest.mark.parametrize("L", [32, 48])
def test_force_larger_sizes(oracle_mod, L):
    p = batio3_zvr(L)
    o, g = _pair(oracle_mod, p)
    u = _random(p, 50 + L)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    _check_energies(eg, eo)


def test_np_64_cubed(oracle_mod):
    p = batio3_zvr(64, kelvin=300.0)
    p.Q_Nose = 1.0                      # 18example-benchmark/md64x64x64.feram:6
    o, g = _pair(oracle_mod, p)
    u = np.random.default_rng(64).uniform(-0.05, 0.05, size=(64, 64, 64, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    for _ in range(4):
        o.nose_poincare(e_o)
    eg = g.nose_poincare(4)
    assert _normwise(g.get("dipoR"), o.field("dipoR")) <= TRAJ_TOL
    assert _normwise(g.get("dipoP"), o.field("dipoP")) <= TRAJ_TOL
    _check_energies(eg, e_o.as_dict(), tol=TRAJ_TOL)
    assert g.nose[0] == pytest.approx(o.nose[0], rel=TRAJ_TOL)


def test_acoustic_vs_48_cubed(oracle_mod):
    p = batio3_zvr(48, method="vs", kelvin=300.0, acoustic_mass_amu=46.44)
    o, g = _pair(oracle_mod, p)
    u = np.random.default_rng(48).uniform(-0.05, 0.05, size=(48, 48, 48, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    for _ in range(3):
        o.velocity_scaling(e_o)
    eg = g.velocity_scaling(3)
    for name in ("dipoR", "dipoP", "acouR", "acouP"):
        assert _normwise(g.get(name), o.field(name)) <= TRAJ_TOL, name
    _check_energies(eg, e_o.as_dict(), tol=TRAJ_TOL)


def test_128_cubed_force_is_gradient():
    """Size-independent check at the benchmark size: central differences of the
    reported potential energy reproduce dV/du (no oracle at this size)."""
    p = batio3_zvr(128)
    g = _system(p)
    u = np.random.default_rng(128).uniform(-0.05, 0.05, size=(128, 128, 128, 3)) + 0.05
    g.put("dipoR", u)
    g.mirror_force_mirror()
    F = g.get("dVddi")
    assert np.all(np.isfinite(F))

    def V():
        e = g.mirror_force_mirror()
        return sum(e[t] for t in ENERGY_TERMS)

    h = 1e-4
    rng = np.random.default_rng(1)
    for _ in range(3):
        idx = (rng.integers(128), rng.integers(128), rng.integers(128), rng.integers(3))
        up, um = u.copy(), u.copy()
        up[idx] += h
        um[idx] -= h
        g.put("dipoR", up)
        vp = V()
        g.put("dipoR", um)
        vm = V()
        fd = (vp - vm) / (2 * h)
        assert fd == pytest.approx(F[idx], rel=1e-4, abs=1e-6 * np.max(np.abs(F)))


# ---- FP32 mode: spectra, FFTs, the k-space stage and the stored state (u, p)
# in FP32; local forces, kicks and drifts computed in FP64 registers;
# thermostat and energy sums in FP64.  Bound from the FP32-mode target:
# force error <= 1e-5 relative (normwise) against the FP64 oracle.
FP32_TOL = 1e-5


def _system32(p):
    from feram_amd.system import System
    print(f"[gpu] System FP32 L=({p.Lx},{p.Ly},{p.Lz}) method={p.method}", flush=True)
    return System(p, precision=1)


# (12, 6, 128) and (6, 81, 64): the paired z pass's last tile is partial (and
# at 6 x 81 its FP32 rows do not start on 16 bytes), so its prefetch falls back
# to direct loads for those pencils
@pytest.mark.parametrize("L", [12, 16, (8, 12, 16), 32, (8, 8, 256), (12, 6, 128), (6, 16, 64), (16, 16, 80),
                               (40, 90, 12), (16, 16, 81), (6, 81, 64)])
def test_fp32_force_and_energies(oracle_mod, L):
    p = batio3_zvr(L)
    o = oracle_mod.Oracle(p)
    g = _system32(p)
    u = _random(p, 70)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    err = _normwise(g.get("dVddi"), o.field("dVddi"))
    print(f"[gpu] FP32 force error {err:.2e}")
    assert err <= FP32_TOL
    _check_energies(eg, eo, tol=FP32_TOL)
    lo = o.field("long_range")
    assert _normwise(g.get_long_range()[:, :p.Ly], lo) <= 1e-12      # the table is built in FP64


def test_fp32_acoustic_force(oracle_mod):
    p = batio3_zvr((8, 12, 8), method="lf", acoustic_mass_amu=46.44)
    o = oracle_mod.Oracle(p)
    g = _system32(p)
    u = _random(p, 71)
    w = 0.02 * np.random.default_rng(72).standard_normal(u.shape)
    for name, v in (("dipoR", u), ("acouR", w)):
        o.field(name)[...] = v
        g.put(name, v)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FP32_TOL
    assert _normwise(g.get("dVdacR"), o.field("dVdacR")) <= FP32_TOL
    _check_energies(eg, eo, tol=FP32_TOL)


def test_fp32_np_trajectory(oracle_mod):
    p = batio3_zvr(16, kelvin=300.0)
    o = oracle_mod.Oracle(p)
    g = _system32(p)
    u = np.random.default_rng(73).uniform(-0.05, 0.05, size=(16, 16, 16, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    for _ in range(20):
        o.nose_poincare(e_o)
    eg = g.nose_poincare(20)
    assert _normwise(g.get("dipoR"), o.field("dipoR")) <= FP32_TOL
    assert _normwise(g.get("dipoP"), o.field("dipoP")) <= FP32_TOL
    _check_energies(eg, e_o.as_dict(), tol=FP32_TOL)
    assert g.nose[0] == pytest.approx(o.nose[0], rel=FP32_TOL)


def test_fp32_state_is_stored_in_fp32():
    """FP32 mode stores u and p as FP32 words: a put rounds to nearest, a get
    widens exactly, and the FP64 mode keeps every bit."""
    p = batio3_zvr((8, 12, 6))
    rng = np.random.default_rng(31)
    u = rng.uniform(-0.1, 0.1, size=(p.Lz, p.Ly, p.Lx, 3))
    pm = rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)) * 1e-3
    g32, g64 = _system32(p), _system(p)
    for g in (g32, g64):
        g.put("dipoR", u)
        g.put("dipoP", pm)
    assert np.array_equal(g32.get("dipoR"), u.astype(np.float32).astype(np.float64))
    assert np.array_equal(g32.get("dipoP"), pm.astype(np.float32).astype(np.float64))
    assert np.array_equal(g64.get("dipoR"), u) and np.array_equal(g64.get("dipoP"), pm)
    g32.close()
    g64.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_set_modulation_leaves_the_state(precision):
    """fhip_set_modulation is a set-up call: dV/du (which the next call's first
    kick reads), u and p stay as they are."""
    from feram_amd.system import System
    g_ = load_golden()["modulation_check"]
    p = param_from_dict(g_["params"])
    mod = np.zeros((p.Lz, p.Ly, p.Lx))
    for x, y, z, a in g_["modulation"]:
        mod[z, y, x] = a
    p.method = "md"                    # Nose-Poincare: its first kick reads the stored dV/du
    g = System(p, precision=precision)
    g.put("dipoR", _random(p, 12))
    g.md_prologue()
    g.nose_poincare(2)
    before = {k: g.get(k) for k in ("dVddi", "dipoR", "dipoP")}
    g.set_modulation(mod)
    for k, v in before.items():
        assert np.array_equal(g.get(k), v), k
    g.close()


def test_fp32_modulation_energies():
    g_ = load_golden()["modulation_check"]
    p = param_from_dict(g_["params"])
    mod = np.zeros((p.Lz, p.Ly, p.Lx))
    for x, y, z, a in g_["modulation"]:
        mod[z, y, x] = a
    g = _system32(p)
    g.set_modulation(mod)
    e = g.md_prologue()
    assert e["inho_strain"] / p.N == pytest.approx(0.000072004971, rel=1e-6)
    assert e["inho_modulation"] / p.N == pytest.approx(-0.000144009942, rel=1e-6)


ZSTAR_SPECIES = [9.956, 7.5, 12.25]


def _species_pair(oracle_mod, p, seed, precision=0):
    """Oracle and device with relaxor species of different Z* (read_relaxor.F:35)."""
    from feram_amd.system import System
    rng = np.random.default_rng(seed)
    relax = rng.integers(1, 4, size=(p.Lz, p.Ly, p.Lx)).astype(np.int32)
    o = oracle_mod.Oracle(p)
    o.relaxor[...] = relax
    o.field("Z_starR")[...] = np.array(p.Z_star)[relax - 1]
    for s in range(3):
        o.field("mass_dipo_inv")[relax == s + 1] = 1.0 / p.mass_dipo[s]
    g = System(p, precision=precision)
    g.set_cells(relaxor=relax)
    u = _random(p, seed + 1)
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    if p.mass_acou[0] > 0:
        w = 0.02 * rng.standard_normal(u.shape)
        o.field("acouR")[...] = w
        g.put("acouR", w)
    return o, g


@pytest.mark.parametrize("L,method,acoustic", [((12, 8, 16), "md", False), ((16, 16, 80), "vs", False),
                                               ((8, 12, 8), "lf", True)])
def test_species_zstar_force_and_trajectory(oracle_mod, L, method, acoustic):
    """Species of different Z*: u and Z*(R) u transformed separately, Z*(R) on
    the long-range force and on the field term, as the reference does."""
    p = batio3_zvr(L, method=method, kelvin=300.0, acoustic_mass_amu=46.44 if acoustic else None)
    p.Z_star = list(ZSTAR_SPECIES)
    p.mass_dipo = [p.mass_dipo[0], 1.3 * p.mass_dipo[0], 0.7 * p.mass_dipo[0]]
    p.external_E_field = [0.0008, -0.0005, 0.0015]
    p.E_field_always_zero = False
    print(f"[gpu] species Z* L={L} {method} acoustic={acoustic}", flush=True)
    o, g = _species_pair(oracle_mod, p, 61)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FORCE_TOL
    _check_energies(eg, eo)
    if acoustic:
        assert _normwise(g.get("dVdacR"), o.field("dVdacR")) <= FORCE_TOL
    e_o = o.md_prologue()
    g.md_prologue()
    for _ in range(6):
        o.step(e_o)
    eg = g.step(6)
    for name in ("dipoR", "dipoP", "dVddi"):
        assert _normwise(g.get(name), o.field(name)) <= TRAJ_TOL, name
    _check_energies(eg, e_o.as_dict(), TRAJ_TOL)
    g.close()


def test_species_zstar_fp32(oracle_mod):
    p = batio3_zvr((16, 12, 16))
    p.Z_star = list(ZSTAR_SPECIES)
    o, g = _species_pair(oracle_mod, p, 62, precision=1)
    eo = o.mirror_force_mirror().as_dict()
    eg = g.mirror_force_mirror()
    assert _normwise(g.get("dVddi"), o.field("dVddi")) <= FP32_TOL
    _check_energies(eg, eo, tol=FP32_TOL)
    g.close()


# The paired z kernel runs the two groups of a tile one after the other in one
# workgroup, or side by side in two (chosen from the grid's size;
# FHIP_ZSPLIT=0/1 forces either); both compute every point the same way.
@pytest.mark.parametrize("L,method,acoustic,precision", [
    (16, "vs", 46.44, 0), (64, "md", None, 0), ((12, 6, 128), "md", None, 0), (64, "md", None, 1)])
def test_z_pair_split_matches_unsplit(monkeypatch, L, method, acoustic, precision):
    from feram_amd.system import System
    p = batio3_zvr(L, method=method, kelvin=300.0, acoustic_mass_amu=acoustic)
    p.Q_Nose = 1.0
    u = _random(p, 7)
    out = []
    for zs in ("0", "1"):
        monkeypatch.setenv("FHIP_ZSPLIT", zs)
        print(f"[gpu] z pair split={zs} L={L} {method} precision={precision}", flush=True)
        g = System(p, precision=precision)
        g.put("dipoR", u)
        g.md_prologue()
        e = (g.velocity_scaling if method == "vs" else g.nose_poincare)(5)
        names = ("dipoR", "dipoP") + (("acouR", "acouP") if acoustic else ())
        out.append(({n: g.get(n) for n in names}, e))
        g.close()
    tol = 1e-12 if precision == 0 else FP32_TOL
    (f0, e0), (f1, e1) = out
    for n in f0:
        assert _normwise(f1[n], f0[n]) <= tol, n
    _check_energies(e1, e0, tol=tol)
