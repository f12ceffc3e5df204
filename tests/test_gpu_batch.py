"""Replica batches (fhip_batch) against the same replicas stepped one by one.

A batch runs the per-context stage bodies with blockIdx.z selecting the
replica, so every replica must evolve as its own step call makes it evolve:
the fields, thermostat state and energies are compared with those of twin
contexts stepped individually.  Replicas differ in temperature and field (the
sweep a batch serves) and fall into several size classes (edges, precision).
The graph replay of the middle steps must give exactly what direct launches
give.  Individual steps are checked against the oracle in the other GPU
tests, so agreement here carries that parity over to the batch.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

# Same arithmetic in the same order; the batched kernels are separate
# instantiations, so the compiler's FMA contraction may differ in the last bit.
TOL64 = 1e-12
TOL32 = 2e-5

SWEEP = [  # (edges, precision, kelvin, E_z in V/A)
    ((8, 8, 8), 0, 150.0, 0.0),
    ((8, 8, 8), 0, 300.0, 0.001),
    ((12, 12, 12), 0, 450.0, 0.0005),
    ((8, 8, 8), 1, 250.0, 0.0),
    ((16, 8, 12), 0, 200.0, 0.0),
    ((20, 40, 60), 0, 350.0, 0.0005),
    ((8, 8, 128), 0, 300.0, 0.0005),
    ((96, 8, 8), 1, 320.0, 0.0),
]


def make(edges, prec, kelvin, ez, method, seed):
    from feram_amd.system import System
    Lx, Ly, Lz = edges
    p = batio3_zvr((Lx, Ly, Lz), method=method, kelvin=kelvin)
    p.external_E_field = [0.0, 0.0, ez]
    p.E_field_always_zero = ez == 0.0
    p.Q_Nose = 0.1 * p.N / 16 ** 3
    g = System(p, precision=prec)
    rng = np.random.default_rng(seed)
    g.put("dipoR", 0.05 + 0.04 * rng.standard_normal((Lz, Ly, Lx, 3)))
    g.put("dipoP", 0.02 * rng.standard_normal((Lz, Ly, Lx, 3)))
    return g


def pair(method):
    a = [make(e, pr, T, ez, method, 100 + i) for i, (e, pr, T, ez) in enumerate(SWEEP)]
    b = [make(e, pr, T, ez, method, 100 + i) for i, (e, pr, T, ez) in enumerate(SWEEP)]
    return a, b


def rel(x, y):
    return np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300)


def compare(single, batched, energies_s, energies_b):
    for i, (s, b) in enumerate(zip(single, batched)):
        tol = TOL32 if SWEEP[i][1] == 1 else TOL64
        same = all(np.array_equal(b.get(f), s.get(f)) for f in ("dipoR", "dipoP", "dVddi"))
        print(f"[gpu] replica {i}: u err {rel(b.get('dipoR'), s.get('dipoR')):.2e}, bitwise {same}", flush=True)
        for f in ("dipoR", "dipoP", "dVddi"):
            assert rel(b.get(f), s.get(f)) <= tol, (i, f)
        ns, nb = np.array(s.nose), np.array(b.nose)
        assert np.max(np.abs(nb - ns)) <= tol * max(np.max(np.abs(ns)), 1.0), i
        assert rel(b.strain, s.strain) <= max(tol, 1e-10), i
        scale = sum(abs(v) for v in energies_s[i].values())
        for k, v in energies_s[i].items():
            assert abs(energies_b[i][k] - v) <= tol * scale, (i, k)


def test_np_batch_matches_individual_steps():
    from feram_amd.system import Batch
    print("[gpu] batch md: 8 replicas, 7 size classes, 22 steps", flush=True)
    single, batched = pair("md")
    es = []
    for s in single:
        s.md_prologue()
    for b in batched:
        b.md_prologue()
    nsteps = 22      # 20 middle steps: the 16- and 4-step graphs
    for s in single:
        es.append(s.nose_poincare(nsteps))
    with Batch(batched) as bt:
        assert bt.classes == 7
        eb = bt.step(nsteps)
    compare(single, batched, es, eb)
    # a second call continues from the state the first left
    es2 = [s.nose_poincare(3) for s in single]
    with Batch(batched) as bt:
        eb2 = bt.step(3)
    compare(single, batched, es2, eb2)


@pytest.mark.parametrize("method", ["lf", "vs"])
def test_lf_vs_batch_matches_individual_steps(method):
    from feram_amd.system import Batch
    print(f"[gpu] batch {method}: 5 replicas, 5 steps", flush=True)
    single, batched = pair(method)
    for s in single:
        s.mirror_force_mirror()
    for b in batched:
        b.mirror_force_mirror()
    es = [s.step(5) for s in single]
    with Batch(batched) as bt:
        eb = bt.step(5)
    compare(single, batched, es, eb)


def test_graph_replay_equals_direct_launches():
    from feram_amd.system import Batch
    print("[gpu] batch graph vs direct", flush=True)
    a, b = pair("md")
    for s in a + b:
        s.md_prologue()
    with Batch(a) as ba, Batch(b) as bb:
        bb.use_graph(False)
        ea = ba.step(6)
        eb = bb.step(6)
        ea2 = ba.step(2)      # graph replayed again by a later call
        eb2 = bb.step(2)
    for x, y in zip(a, b):
        for f in ("dipoR", "dipoP"):
            assert np.array_equal(x.get(f), y.get(f)), f
    assert ea == eb and ea2 == eb2


def test_batch_averages_match_individual():
    from feram_amd.system import Batch
    print("[gpu] batch averages", flush=True)
    single, batched = pair("md")
    for g in single + batched:
        g.md_prologue()
        g.avg_reset(histograms=False, dipoRavg=True)
        g.avg_auto(True)
    for s in single:
        s.nose_poincare(4)
    with Batch(batched) as bt:
        bt.step(4)
    for i, (s, b) in enumerate(zip(single, batched)):
        tol = TOL32 if SWEEP[i][1] == 1 else TOL64
        a_s, a_b = s.avg_get(), b.avg_get()
        assert a_b["n_add"] == a_s["n_add"] == 4
        for key in ("u", "uu", "p", "pp", "strain"):
            x, y = np.array(a_b[key]), np.array(a_s[key])
            assert np.max(np.abs(x - y)) <= tol * max(np.max(np.abs(y)), 1e-30), (i, key)
        assert rel(b.get("dipoRavg"), s.get("dipoRavg")) <= tol, i


def test_batch_rejects_what_it_cannot_run():
    from feram_amd.system import Batch, FhipError
    print("[gpu] batch argument checks", flush=True)
    a = make((8, 8, 8), 0, 300.0, 0.0, "md", 1)
    b = make((8, 8, 8), 0, 300.0, 0.0, "lf", 2)
    with pytest.raises(FhipError, match="same method"):
        Batch([a, b])
    with pytest.raises(FhipError, match="twice"):
        Batch([a, a])
    c = make((8, 8, 8), 0, 300.0, 0.0, "md", 4)
    a.avg_reset()
    a.avg_auto(True)
    with Batch([a, c]) as bt:
        with pytest.raises(FhipError, match="averaging settings"):
            bt.step(1)


def test_autotuned_plan_agrees_with_untuned():
    """The tuner picks, per size class, the full or half tile for the x rows
    and the y lines; either plan computes the same step to rounding."""
    from feram_amd.system import Batch
    print("[gpu] batch tile-shape tuner", flush=True)
    a, b = pair("md")
    for s in a + b:
        s.md_prologue()
    with Batch(a) as ba, Batch(b) as bb:
        bb.tune(False)
        assert not any(p["tuned"] for p in ba.plan())
        ea = ba.step(6)
        eb = bb.step(6)
        plan, fixed = ba.plan(), bb.plan()
    for p, q in zip(plan, fixed):
        assert p["tuned"] and not q["tuned"]
        assert p["L"] == q["L"]
        assert p["x_rows"] in (q["x_rows"], 2 * q["x_rows"])
        assert p["y_lines"] in (q["y_lines"], q["y_lines"] // 2, 1)
        print(f"  class {p['L']}: x rows {p['x_rows']} (us {p['x_us'][0]:.1f}/{p['x_us'][1]:.1f}), "
              f"y lines {p['y_lines']} (us {p['y_us'][0]:.1f}/{p['y_us'][1]:.1f})", flush=True)
    compare(a, b, ea, eb)


@pytest.mark.parametrize("method,acoustic", [("md", False), ("lf", True)])
def test_species_zstar_replicas_in_a_batch(method, acoustic):
    """Replicas whose relaxor species carry different Z* (their own size class:
    Z*(R) u transformed separately, the split dipole-dipole z group) next to
    replicas of the same edges with one Z*, in one batch, against twin
    contexts stepped individually (9-step calls: 4 + 4 middle steps)."""
    from feram_amd.system import Batch, System
    print(f"[gpu] batch with species Z*: {method} acoustic={acoustic}", flush=True)
    specs = [((12, 8, 16), [9.956, 7.5, 12.25], 300.0), ((12, 8, 16), [9.956, 7.5, 12.25], 350.0),
             ((12, 8, 16), [9.956, 9.956, 9.956], 300.0), ((8, 8, 8), [9.956, 8.0, 11.0], 250.0)]

    def make_z(i, edges, zstar, kelvin):
        p = batio3_zvr(edges, method=method, kelvin=kelvin, acoustic_mass_amu=46.44 if acoustic else None)
        p.Z_star = list(zstar)
        p.mass_dipo = [p.mass_dipo[0], 1.3 * p.mass_dipo[0], 0.7 * p.mass_dipo[0]]
        p.external_E_field = [0.0008, -0.0005, 0.0015]
        p.E_field_always_zero = False
        p.Q_Nose = 0.1 * p.N / 16 ** 3
        g = System(p)
        rng = np.random.default_rng(70 + i)
        shape = (p.Lz, p.Ly, p.Lx)
        g.set_cells(relaxor=rng.integers(1, 4, size=shape).astype(np.int32))
        g.put("dipoR", 0.05 + 0.04 * rng.standard_normal(shape + (3,)))
        g.put("dipoP", 0.02 * rng.standard_normal(shape + (3,)))
        if acoustic:
            g.put("acouR", 0.02 * rng.standard_normal(shape + (3,)))
        if method == "md":
            g.md_prologue()
        else:
            g.mirror_force_mirror()
        return g

    single = [make_z(i, *sp) for i, sp in enumerate(specs)]
    batched = [make_z(i, *sp) for i, sp in enumerate(specs)]
    try:
        es = [s.step(9) for s in single]
        with Batch(batched) as bt:
            assert bt.classes == 3
            eb = bt.step(9)
        for i, (s, b) in enumerate(zip(single, batched)):
            for f in ("dipoR", "dipoP", "dVddi") + (("acouR", "acouP") if acoustic else ()):
                assert rel(b.get(f), s.get(f)) <= TOL64, (i, f)
            scale = sum(abs(v) for v in es[i].values())
            for k, v in es[i].items():
                assert abs(eb[i][k] - v) <= TOL64 * scale, (i, k)
    finally:
        for g in single + batched:
            g.close()
