"""The 128^3 NP-MD workload against the FP64 oracle, FP64 plain and FP32 mode.

BaTiO3 bulk 128x128x128 (2,097,152 cells), 300 K, dt = 2 fs, Q_Nose 8.0
(the reference's 64^3 md benchmark value, 1.0, scaled with N), u0 uniform in
[-0.05, 0.05] A from a fixed seed, p = 0: the prologue's force evaluation
(initialize-dipoR.F:59-69) and two Nose-Poincare steps (nose-poincare.F:10-78)
on the device and in the oracle from the same doubles.

Tolerances (normwise max|d| / max|ref| for fields, energy terms against the
sum of |terms|):
  FP64 plain mode: forces 1e-12 after the prologue; u, p, forces 1e-10 and
                   energies 1e-10 after two steps; s_Nose, pi_Nose 1e-10.
  FP32 mode:       forces and energies 1e-5, u and p after two steps 1e-5.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

ENERGY_TERMS = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain",
                "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation")
L = 128


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _energy_err(eg, eo):
    scale = sum(abs(eo[t]) for t in ENERGY_TERMS)
    return max(abs(eg[t] - eo[t]) for t in ENERGY_TERMS) / scale


@pytest.fixture(scope="module")
def reference(oracle_mod):
    p = batio3_zvr(L, method="md", kelvin=300.0)
    p.Q_Nose = 1.0 * p.N / 64 ** 3
    rng = np.random.default_rng(20241203)
    u0 = rng.uniform(-0.05, 0.05, size=(L, L, L, 3))
    print("[gpu] 128^3 oracle: set-up, prologue and 2 NP steps", flush=True)
    o = oracle_mod.Oracle(p)
    o.field("dipoR")[...] = u0
    e = o.md_prologue()
    out = {"p": p, "u0": u0, "F0": o.field("dVddi").copy(), "e0": e.as_dict()}
    for _ in range(2):
        o.nose_poincare(e)
    out.update(u2=o.field("dipoR").copy(), p2=o.field("dipoP").copy(), F2=o.field("dVddi").copy(),
               e2=e.as_dict(), nose=tuple(o.nose))
    o.close()
    return out


@pytest.mark.parametrize("precision,tol0,tol", [(0, 1e-12, 1e-10), (1, 1e-5, 1e-5)], ids=["fp64", "fp32"])
def test_md128_against_oracle(reference, precision, tol0, tol):
    from feram_amd.system import System
    r = reference
    print(f"[gpu] 128^3 NP-MD precision={precision}: prologue + 2 steps", flush=True)
    with System(r["p"], precision=precision) as g:
        g.put("dipoR", r["u0"])
        e0 = g.md_prologue()
        F0_err = _rel(g.get("dVddi"), r["F0"])
        E0_err = _energy_err(e0, r["e0"])
        e2 = g.nose_poincare(2)
        errs = {"u": _rel(g.get("dipoR"), r["u2"]), "p": _rel(g.get("dipoP"), r["p2"]),
                "F": _rel(g.get("dVddi"), r["F2"]), "E": _energy_err(e2, r["e2"])}
        s, pi, _ = g.nose
    print(f"[gpu]   prologue: force {F0_err:.2e} energy {E0_err:.2e}; after 2 steps: "
          + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert F0_err <= tol0 and E0_err <= tol0
    for k, v in errs.items():
        assert v <= tol, (k, v)
    so, pio, _ = r["nose"]
    assert abs(s - so) <= tol * abs(so)
    assert abs(pi - pio) <= tol * max(abs(pio), 1e-300) + 1e-14
