"""Leapfrog with heat-source and sink layers (leapfrog.F:38-83) against the oracle.

A 10-step leapfrog trajectory with delta_e (and, with acoustic MD,
delta_e_acou) on an 8x6x16 supercell with a sum_p mask: positions, momenta
and the layer energies of every step must follow the oracle (pinned in
tests/test_oracle_heat_source.py) to the trajectory tolerance, and the
averaged layer sums must equal the sum of the per-step values.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-10


@pytest.mark.parametrize("acoustic", [False, True])
def test_leapfrog_heat_source_matches_oracle(oracle_mod, acoustic):
    from feram_amd.system import System
    print(f"[gpu] heat-source layers, acoustic={acoustic}", flush=True)
    p = batio3_zvr((8, 6, 16), method="lf", kelvin=300.0, acoustic_mass_amu=46.44 if acoustic else None)
    o = oracle_mod.Oracle(p)
    g = System(p)
    rng = np.random.default_rng(9)
    u = 0.06 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    pm = 0.1 * rng.standard_normal(u.shape)
    sum_p = np.ones((p.Lz, p.Ly, p.Lx), np.int32)
    sum_p[4, 0, :3] = 0
    o.sum_p[...] = sum_p.astype(np.uint8)
    o.recount_sum_p()
    g.set_cells(sum_p=sum_p)
    for name, v in (("dipoR", u), ("dipoP", pm)):
        o.field(name)[...] = v
        g.put(name, v)
    if acoustic:
        pw = 0.05 * rng.standard_normal(u.shape)
        o.field("acouP")[...] = pw
        g.put("acouP", pw)
    de, dea = 0.002, (0.001 if acoustic else 0.0)
    o.set_heat_source(de, dea)
    g.set_heat_source(de, dea)
    e = o.mirror_force_mirror()
    g.mirror_force_mirror()
    g.avg_reset()
    g.avg_auto(True)
    sums = np.zeros((2, p.Lz))
    for step in range(10):
        o.leapfrog(e, 1)
        g.leapfrog(1, 1)
        ed, ea = g.layer_energies()
        od, oa = o.layers()
        sums += (ed, ea)
        assert np.allclose(ed, od, rtol=1e-11, atol=0), step
        assert np.allclose(ea, oa, rtol=1e-11, atol=1e-300), step
    for name in ("dipoR", "dipoP") + (("acouR", "acouP") if acoustic else ()):
        a, b = g.get(name), o.field(name)
        assert np.max(np.abs(a - b)) <= TRAJ_TOL * np.max(np.abs(b)), name
    sd, sa = g.layer_energies(averaged=True)
    assert np.allclose(sd, sums[0], rtol=1e-12) and np.allclose(sa, sums[1], rtol=1e-12, atol=1e-300)
    g.close()


def test_leapfrog_heat_source_fp32(oracle_mod):
    """FP32 mode (FP32 spectra and stored state) with FP64 layer sums, 10 steps."""
    from feram_amd.system import System
    print("[gpu] heat-source layers, fp32", flush=True)
    p = batio3_zvr((16, 16, 16), method="lf", kelvin=300.0)
    o = oracle_mod.Oracle(p)
    g = System(p, precision=1)
    rng = np.random.default_rng(19)
    u = 0.06 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    pm = 0.1 * rng.standard_normal(u.shape)
    for name, v in (("dipoR", u), ("dipoP", pm)):
        o.field(name)[...] = v
        g.put(name, v)
    o.set_heat_source(0.002)
    g.set_heat_source(0.002)
    e = o.mirror_force_mirror()
    g.mirror_force_mirror()
    for _ in range(10):
        o.leapfrog(e, 1)
    g.leapfrog(1, 10)
    ed, _ = g.layer_energies()
    od, _ = o.layers()
    assert np.allclose(ed, od, rtol=2e-5, atol=0)
    for name in ("dipoR", "dipoP"):
        a, b = g.get(name), o.field(name)
        assert np.max(np.abs(a - b)) <= 2e-5 * np.max(np.abs(b)), name
    g.close()


def test_heat_source_in_a_batch_matches_individual_steps():
    """Leapfrog replicas with and without heat-source layers (one with acoustic
    MD and delta_e_acou, one with delta_e only, one without) stepped as one
    batch, against twin contexts stepped individually: fields, layer energies
    and their averaged sums after 12 steps (16-step graphs are not reached; 4
    and 1 are) and after a second call of 20 steps (16 + 1 + 1 + 1)."""
    from feram_amd.system import Batch, System
    print("[gpu] heat-source layers in a batch", flush=True)
    cases = [((8, 6, 16), True, 0.002, 0.001), ((8, 6, 16), False, 0.003, 0.0), ((12, 12, 12), False, 0.0, 0.0),
             ((8, 6, 16), False, 0.001, 0.0)]

    def make(i, edges, acoustic, de, dea):
        p = batio3_zvr(edges, method="lf", kelvin=300.0, acoustic_mass_amu=46.44 if acoustic else None)
        g = System(p)
        rng = np.random.default_rng(40 + i)
        shape = (p.Lz, p.Ly, p.Lx, 3)
        g.put("dipoR", 0.06 + 0.05 * rng.standard_normal(shape))
        g.put("dipoP", 0.1 * rng.standard_normal(shape))
        if acoustic:
            g.put("acouP", 0.05 * rng.standard_normal(shape))
        if de or dea:
            g.set_heat_source(de, dea)
        g.mirror_force_mirror()
        g.avg_reset()
        g.avg_auto(True)
        return g

    single = [make(i, *c) for i, c in enumerate(cases)]
    batched = [make(i, *c) for i, c in enumerate(cases)]
    try:
        for nsteps in (12, 20):
            es = [s.leapfrog(1, nsteps) for s in single]
            with Batch(batched) as bt:
                eb = bt.step(nsteps)
            for i, (s, b) in enumerate(zip(single, batched)):
                for f in ("dipoR", "dipoP") + (("acouR", "acouP") if cases[i][1] else ()):
                    x, y = b.get(f), s.get(f)
                    assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(y)), (nsteps, i, f)
                scale = sum(abs(v) for v in es[i].values())
                for k, v in es[i].items():
                    assert abs(eb[i][k] - v) <= 1e-12 * scale, (nsteps, i, k)
                if cases[i][2] or cases[i][3]:
                    for avg in (False, True):
                        xs, ys = b.layer_energies(averaged=avg), s.layer_energies(averaged=avg)
                        for x, y in zip(xs, ys):
                            assert np.allclose(x, y, rtol=1e-12, atol=1e-300), (nsteps, i, avg)
    finally:
        for g in single + batched:
            g.close()
