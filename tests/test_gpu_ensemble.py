"""The 48-replica batched ensemble, FP64 and FP32 mode.

The replicas of feram_amd.sweeps.replica_ensemble: BaTiO3 bulk NP-MD, edges
drawn from {12, 16, 24, 32} (an uneven mix, four size classes), T uniform in
150-450 K and a static E_z uniform in 0-100 kV/cm, u0 uniform in
[-0.05, 0.05] A.  One batch steps all 48 (per-class streams, the captured
graph and the autotuned tile plans); every replica is compared with a twin
context stepped on its own, and the first replica of every size class with
the FP64 oracle.

Tolerances (normwise max|d| / max|ref|; energies against the sum of |terms|):
  batch vs twin:    FP64 1e-12, FP32 2e-5 (separate kernel instantiations,
                    so FMA contraction may differ in the last bit)
  batch vs oracle:  FP64 1e-10, FP32 1e-5 after 5 steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 5
ENERGY_TERMS = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain",
                "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _energy_err(eg, eo):
    scale = sum(abs(eo[t]) for t in ENERGY_TERMS)
    return max(abs(eg[t] - eo[t]) for t in ENERGY_TERMS) / scale


@pytest.fixture(scope="module")
def ensemble():
    from feram_amd.sweeps import replica_ensemble
    return replica_ensemble(48)


@pytest.fixture(scope="module")
def oracle_runs(oracle_mod, ensemble):
    """Oracle state after the prologue and STEPS NP steps of the first replica
    of every edge."""
    params, u0 = ensemble
    firsts = {}
    for i, p in enumerate(params):
        firsts.setdefault(p.Lx, i)
    out = {}
    for L, i in sorted(firsts.items()):
        print(f"[gpu] oracle replica {i} ({L}^3, {params[i].kelvin:.0f} K, "
              f"E_z {params[i].external_E_field[2]:.2e})", flush=True)
        o = oracle_mod.Oracle(params[i])
        o.field("dipoR")[...] = u0[i]
        e = o.md_prologue()
        for _ in range(STEPS):
            o.nose_poincare(e)
        out[i] = {"u": o.field("dipoR").copy(), "p": o.field("dipoP").copy(), "F": o.field("dVddi").copy(),
                  "e": e.as_dict(), "nose": tuple(o.nose)}
        o.close()
    return out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_ensemble_batch(ensemble, oracle_runs, precision):
    from feram_amd.system import Batch, System
    params, u0 = ensemble
    tol_twin = 2e-5 if precision else 1e-12
    tol_oracle = 1e-5 if precision else 1e-10
    print(f"[gpu] 48-replica ensemble precision={precision}: batch and twins, {STEPS} steps", flush=True)
    batched, twins = [], []
    for p, u in zip(params, u0):
        for group in (batched, twins):
            g = System(p, precision=precision)
            g.put("dipoR", u)
            g.md_prologue()
            group.append(g)
    try:
        with Batch(batched) as bt:
            assert bt.classes == 4
            eb = bt.step(STEPS)
            plan = bt.plan()
        assert all(c["tuned"] for c in plan)
        et = [t.nose_poincare(STEPS) for t in twins]
        worst = 0.0
        for i, (b, t) in enumerate(zip(batched, twins)):
            for f in ("dipoR", "dipoP", "dVddi"):
                err = _rel(b.get(f), t.get(f))
                worst = max(worst, err)
                assert err <= tol_twin, (i, f, err)
            assert _energy_err(eb[i], et[i]) <= tol_twin, i
            assert abs(b.nose[0] - t.nose[0]) <= tol_twin * abs(t.nose[0]), i
        print(f"[gpu]   batch vs twins: worst field error {worst:.2e}", flush=True)
        for i, r in oracle_runs.items():
            g = batched[i]
            errs = {"u": _rel(g.get("dipoR"), r["u"]), "p": _rel(g.get("dipoP"), r["p"]),
                    "F": _rel(g.get("dVddi"), r["F"]), "E": _energy_err(eb[i], r["e"])}
            print(f"[gpu]   replica {i} ({params[i].Lx}^3) vs oracle: "
                  + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
            for k, v in errs.items():
                assert v <= tol_oracle, (i, k, v)
            assert abs(g.nose[0] - r["nose"][0]) <= tol_oracle * abs(r["nose"][0]), i
    finally:
        for g in batched + twins:
            g.close()
