"""The device field schedule (fhip_set_E_schedule) against the oracle.

Every schedule type (wave_module.F's triangular and ramping waves, the
hysteresis-loop sweep) must put, before each step, the field of that step's
index on the device (checked against fo_wave, which the reference's 360.hl
output and closed forms pin); an NP trajectory under a triangular wave must
follow the oracle stepped with the same field set per step as
molecular-dynamics.F:66-76 does; a batch must carry the schedule per
replica; and a shortened run of the reference's 32^3 hysteresis-loop example
(22example-BaTiO3-new-param-hl/360.feram) must reach its saturated
polarisation at both ends of the sweep and switch in between.
"""
import json
import os

import numpy as np
import pytest

from feram_amd.params import Param, batio3_zvr

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

SCHEDULES = [("triangular_cos", 8, 3, 0), ("triangular_sin", 8, 3, 0), ("ramping_on", 4, 6, 0),
             ("ramping_off", 4, 6, 0), ("hysteresis", 0, 3, 10)]


def field_system(method="md", L=8):
    from feram_amd.system import System
    p = batio3_zvr(L, method=method, kelvin=300.0)
    p.E_field_always_zero = False
    return p, System(p)


@pytest.mark.parametrize("kind,period,ntherm,navg", SCHEDULES)
def test_schedule_follows_step_index(oracle_mod, kind, period, ntherm, navg):
    print(f"[gpu] E schedule {kind}", flush=True)
    p, g = field_system()
    rng = np.random.default_rng(1)
    g.put("dipoR", 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    g.md_prologue()
    A = [0.0005, -0.001, 0.002]
    g.set_E_schedule(kind, A, period, ntherm, navg, i_step=1)
    for k in range(12):
        i, E = g.step_info()
        assert i == 1 + k
        Eo = oracle_mod.wave(kind, A, period, ntherm, navg, i)
        assert np.allclose(E, Eo, rtol=1e-15, atol=0), (kind, i, E, Eo)
        g.nose_poincare(1)
    g.close()


def test_np_trajectory_under_triangular_wave(oracle_mod):
    print("[gpu] NP trajectory under a triangular wave, 12^3", flush=True)
    p, g = field_system(L=12)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(5)
    u = 0.08 + 0.06 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    A, period, ntherm = [0.0, 0.001, 0.003], 8, 2
    # step 0 (the prologue) uses wave(i = 0) (initialize_E_field.F:58-66)
    o.field("E_field")[...] = oracle_mod.wave("triangular_cos", A, period, ntherm, 0, 0)
    g.set_E_schedule("triangular_cos", A, period, ntherm, 0, i_step=0)
    e_o = o.md_prologue()
    g.md_prologue()
    g.set_E_schedule("triangular_cos", A, period, ntherm, 0, i_step=1)
    nsteps = 10
    for i in range(1, nsteps + 1):
        o.field("E_field")[...] = oracle_mod.wave("triangular_cos", A, period, ntherm, 0, i)
        o.nose_poincare(e_o)
    eg = g.nose_poincare(nsteps)
    err = np.max(np.abs(g.get("dipoR") - o.field("dipoR"))) / np.max(np.abs(o.field("dipoR")))
    assert err < 1e-10, err
    assert eg["dipole_E_field"] == pytest.approx(e_o.dipole_E_field, rel=1e-9, abs=1e-12)
    g.close()


def test_batch_carries_each_replicas_schedule():
    from feram_amd.system import Batch
    print("[gpu] E schedules in a batch", flush=True)
    pairs = []
    for rep in range(2):
        pair = []
        for kind, period, ntherm, navg in SCHEDULES[:3]:
            p, g = field_system()
            g.put("dipoR", 0.05 * np.random.default_rng(7).standard_normal((p.Lz, p.Ly, p.Lx, 3)))
            g.md_prologue()
            g.set_E_schedule(kind, [0.0, 0.0, 0.002], period, ntherm, navg, i_step=1)
            pair.append(g)
        pairs.append(pair)
    for g in pairs[0]:
        g.nose_poincare(9)
    with Batch(pairs[1]) as bt:
        bt.step(9)
    for a, b in zip(*pairs):
        ia, Ea = a.step_info()
        ib, Eb = b.step_info()
        assert ia == ib == 10 and np.array_equal(Ea, Eb)
        assert np.max(np.abs(a.get("dipoR") - b.get("dipoR"))) <= 1e-12 * np.max(np.abs(a.get("dipoR")))


def test_hysteresis_loop_example_shortened():
    from feram_amd.sweeps import hysteresis_loop
    from feram_amd.system import System
    d = json.load(open(os.path.join(HERE, "golden", "bto32_hysteresis_360K.json")))
    prm = dict(d["param"])
    prm["method"] = "vs"            # 'hl' steps with velocity scaling (hysteresis_loop.F:69)
    p = Param(**prm)
    ntherm, navg, freq = 4000, 100000, 10000
    print(f"[gpu] hysteresis loop 32^3 at 360 K, {ntherm} + {navg} steps (the example sweeps 2000000)",
          flush=True)
    g = System(p)
    rng = np.random.default_rng(11)
    g.put("dipoR", np.array(p.init_dipo_avg) + np.array(p.init_dipo_dev) * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    g.set_E_schedule("hysteresis", p.external_E_field, 0, ntherm, navg, i_step=0)
    g.mirror_force_mirror()

    def progress(step, last):
        r = last[0]
        print(f"[gpu]   step {step}: E_z {r[4]:+.6f}  u_z {r[13]:+.4f}", flush=True)

    rows = hysteresis_loop([g], ntherm, navg, freq, p.external_E_field, i_start=1, progress=progress)[0]
    ref = d["rows"]
    assert [r[0] for r in rows][0] == ntherm and rows[-1][0] == ntherm + navg
    # the field printed with each row is the sweep's (the .hl columns)
    for r in rows:
        assert r[4] == pytest.approx(p.external_E_field[2] * (1 - 2 * (r[0] - ntherm) / navg), abs=1e-15)
    # saturated polarisation at both ends, as the reference's 2000000-step sweep
    assert rows[0][13] == pytest.approx(ref[0][13], abs=0.01)
    assert rows[-1][13] == pytest.approx(ref[-1][13], abs=0.01)
    assert rows[0][13] < -0.1 < 0.1 < rows[-1][13]
    g.close()
