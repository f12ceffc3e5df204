"""Device averages (add_to_Average, average_module.F:57-163) against the oracle.

An NP trajectory at 12^3 with relaxor species and a sum_p mask: after every
step the device accumulates (fhip_avg_auto) and the oracle calls
add_to_Average; the sums agree to the trajectory tolerance and the histograms
exactly (integer counts; no dipole component sits within 1e-10 of a bin edge).
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_average_over_np_trajectory(oracle_mod):
    from feram_amd.system import System
    p = batio3_zvr(12, method="md", kelvin=300.0)
    print("[gpu] averages 12^3 md", flush=True)
    o = oracle_mod.Oracle(p)
    g = System(p)
    rng = np.random.default_rng(41)
    relax = rng.integers(1, 3, size=(p.Lz, p.Ly, p.Lx)).astype(np.int32)
    sum_p = (rng.random((p.Lz, p.Ly, p.Lx)) < 0.8).astype(np.int32)
    o.relaxor[...] = relax
    o.sum_p[...] = sum_p.astype(np.uint8)
    o.recount_sum_p()
    g.set_cells(relaxor=relax, sum_p=sum_p)
    u = 0.08 + 0.06 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    acc = oracle_mod.FoAverage()
    d1 = np.zeros((199, 3), np.int64)
    d3 = np.zeros((199, 199, 199), np.int32)
    g.avg_reset(histograms=True, dipoRavg=True)
    g.avg_auto(True)
    nsteps = 6
    g.nose_poincare(nsteps)
    uavg = np.zeros_like(u)
    for _ in range(nsteps):
        o.nose_poincare(e_o)
        # the driver's total after each step (molecular-dynamics.F:136)
        e_o.total = (e_o.dipo_kinetic / o.nose[0] ** 2 + e_o.acou_kinetic
                     + oracle_mod.lib().fo_sum_potential_energies(e_o))
        o.add_to_average(e_o, acc, d1, d3)
        uavg += o.field("dipoR")           # molecular-dynamics.F:144-148
    ag = g.avg_get(histograms=True)
    ao = acc.as_dict()
    assert ag["n_add"] == nsteps
    for key in ("u", "uu", "p", "pp", "strain", "u_tmp", "uu_tmp", "p_tmp", "pp_tmp"):
        a, b = np.array(ag[key]), np.array(ao[key])
        assert np.max(np.abs(a - b)) <= TOL * max(np.max(np.abs(b)), 1e-30), key
    scale = sum(abs(v) for v in ao["energy"].values())
    for t, v in ao["energy"].items():
        assert abs(ag["energy"][t] - v) <= TOL * scale, t
    assert ag["e2"] == pytest.approx(ao["e2"], rel=TOL)
    assert ag["dipo_kinetic_true"] == pytest.approx(ao["dipo_kinetic_true"], rel=TOL)
    edge = np.abs(np.abs(o.field("dipoR") * 200 - np.floor(o.field("dipoR") * 200)) - 0.5)
    assert edge.min() > 1e-10
    assert np.array_equal(ag["distribution"], d1.T)
    assert np.array_equal(ag["distribution3d"], d3)
    assert np.max(np.abs(g.get("dipoRavg") - uavg)) <= TOL * np.max(np.abs(uavg))
    # an explicit add after the run equals one more oracle call
    g.avg_auto(False)
    g.avg_add()
    o.add_to_average(e_o, acc, d1, d3)
    assert g.avg_get()["n_add"] == nsteps + 1
    assert np.array(g.avg_get()["u"]) == pytest.approx(np.array(acc.as_dict()["u"]), rel=TOL)


def test_avg_record_from_device_sums(oracle_mod, tmp_path):
    """write_avg on a device run gives the record the oracle's add_to_Average
    sums give (to the record's printed digits)."""
    from feram_amd import formats
    from feram_amd.system import System
    p = batio3_zvr(12, method="md", kelvin=300.0)
    print("[gpu] .avg record 12^3 md", flush=True)
    o = oracle_mod.Oracle(p)
    g = System(p)
    rng = np.random.default_rng(7)
    u = 0.1 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    o.field("dipoR")[...] = u
    g.put("dipoR", u)
    e_o = o.md_prologue()
    g.md_prologue()
    acc = oracle_mod.FoAverage()
    d1 = np.zeros((199, 3), np.int64)
    d3 = np.zeros((199, 199, 199), np.int32)
    g.avg_reset()
    g.avg_auto(True)
    nsteps = 4
    g.nose_poincare(nsteps)
    for _ in range(nsteps):
        o.nose_poincare(e_o)
        e_o.total = (e_o.dipo_kinetic / o.nose[0] ** 2 + e_o.acou_kinetic
                     + oracle_mod.lib().fo_sum_potential_energies(e_o))
        o.add_to_average(e_o, acc, d1, d3)
    path = tmp_path / "run.avg"
    line = formats.write_avg(g, str(path))
    assert path.read_text() == line + "\n"
    got = np.array(line.split(), dtype=float)
    want = np.array(formats.avg_values(acc.as_dict() | {"n_add": nsteps}, p), dtype=float)
    assert got.size == 43 and got[0] == 300.0
    # each E field prints a mantissa 0.ddddd (5 or 7 digits): rounding is at
    # most 5e-5 of the value
    assert np.allclose(got, want, rtol=6e-5, atol=1e-12)
    g.close()
