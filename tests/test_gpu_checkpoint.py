"""Exact resume: k steps + save + load into a fresh context + (n - k) steps is
bitwise the state of the context that saved and went on for n - k steps.

The reference resumes from write_system's .coord text (coord_module.F:472-483;
initialize-dipoR.F:22-38), which keeps u to 6 decimals; the binary checkpoint
keeps every double the step reads, so the two runs must agree exactly:
fields, thermostat, strain, energies, averages, histograms and layer
energies.  Cases: NP-MD with averages, histograms and dipoRavg; acoustic
leapfrog with heat-source layers; velocity scaling in FP32 mode; an epit film
under a triangular field wave (the schedule's step index is in the state).
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu


def _make(case):
    from feram_amd.system import System
    if case == "md_avg":
        p = batio3_zvr(12, method="md", kelvin=280.0)
        prec = 0
    elif case == "lf_acoustic_heat":
        p = batio3_zvr((8, 8, 16), method="lf", kelvin=300.0, acoustic_mass_amu=46.44)
        prec = 0
    elif case == "vs_fp32":
        p = batio3_zvr(16, method="vs", kelvin=250.0)
        prec = 1
    else:                                   # epit film under a field wave
        p = batio3_zvr((8, 8, 16), method="md", kelvin=200.0)
        p.bulk_or_film, p.epi_strain = "epit", -0.005
        p.E_field_always_zero = False
    p.Q_Nose = 0.1 * p.N / 16 ** 3
    g = System(p, precision=prec if case != "epit_wave" else 0)
    rng = np.random.default_rng(91)
    g.put("dipoR", 0.05 + 0.04 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    if p.mass_acou[0] > 0:
        g.put("acouR", 0.01 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    if case == "lf_acoustic_heat":
        # the layer rescale needs kinetic energy in the layers (leapfrog.F:44-87)
        g.put("dipoP", 0.3 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
        g.put("acouP", 0.3 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
        g.set_heat_source(0.002, 0.001)
    if case == "epit_wave":
        g.set_E_schedule("triangular_cos", [0.0, 0.0, 0.003], 12, 3, 0, i_step=0)
    return g


def _state(g, acoustic):
    out = {f: g.get(f) for f in ("dipoR", "dipoP", "dVddi")}
    if acoustic:
        for f in ("acouR", "acouP", "dVdacR"):
            out[f] = g.get(f)
    out["nose"] = np.array(g.nose)
    out["strain"] = g.strain
    out["step"] = g.step_info()[0]
    return out


@pytest.mark.parametrize("case", ["md_avg", "lf_acoustic_heat", "vs_fp32", "epit_wave"])
def test_checkpoint_resume_is_bitwise(case, tmp_path):
    from feram_amd.system import System
    print(f"[gpu] checkpoint {case}", flush=True)
    a = _make(case)
    acoustic = a.p.mass_acou[0] > 0
    if case == "epit_wave":
        a.md_prologue()
        a.set_E_schedule("triangular_cos", [0.0, 0.0, 0.003], 12, 3, 0, i_step=1)
    elif a.p.method == "md":
        a.md_prologue()
    else:
        a.mirror_force_mirror()
    if case == "md_avg":
        a.avg_reset(histograms=True, dipoRavg=True)
        a.avg_auto(True)
    a.step(5)
    path = tmp_path / "state.fhipckpt"
    a.save_checkpoint(path)
    ea = a.step(7)
    sa = _state(a, acoustic)

    c = _make(case)
    c.load_checkpoint(path)
    ec = c.step(7)
    sc = _state(c, acoustic)
    diff = {k: (ea[k], ec[k]) for k in ea if ea[k] != ec[k]}
    assert not diff, diff
    for k, v in sa.items():
        assert np.array_equal(np.asarray(sc[k]), np.asarray(v)), k
    if case == "md_avg":
        xa, xc = a.avg_get(histograms=True), c.avg_get(histograms=True)
        assert xa["n_add"] == xc["n_add"] == 12
        for k in ("u", "uu", "p", "pp", "strain", "e2"):
            assert np.array_equal(np.asarray(xa[k]), np.asarray(xc[k])), k
        assert np.array_equal(xa["distribution"], xc["distribution"])
        assert np.array_equal(xa["distribution3d"], xc["distribution3d"])
        assert np.array_equal(a.get("dipoRavg"), c.get("dipoRavg"))
    if case == "lf_acoustic_heat":
        for avg in (False, True):
            for x, y in zip(a.layer_energies(avg), c.layer_energies(avg)):
                assert np.array_equal(x, y)
    a.close()
    c.close()


def test_checkpoint_rejects_another_supercell():
    from feram_amd.system import FhipError, System
    a = System(batio3_zvr(8, method="md"))
    a.md_prologue()
    blob = a.checkpoint()
    b = System(batio3_zvr((8, 8, 12), method="md"))
    with pytest.raises(FhipError, match="another supercell"):
        b.restore(blob)
    with pytest.raises(FhipError, match="not a checkpoint"):
        a.restore(b"x" * len(blob))
    a.close()
    b.close()
