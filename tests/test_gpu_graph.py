"""Multi-step calls replayed from a captured hipGraph (fhip_use_graph) against
the same calls with every kernel launched directly.

The graphs hold 16, 4 and 1 middle steps (the same kernels with the same
arguments, replayed largest first: a 22-step call replays 16 + 4 + 1), so
both runs must agree bit for bit after every call: several calls in a row,
with host work between them (a temperature change, which changes the
captured arguments and must lead to a new capture; reading fields back).
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

CASES = [("md", "bulk", 0.0), ("lf", "bulk", 0.002), ("vs", "epit", 0.0), ("lf", "film", 0.0)]


def _system(method, geometry, delta_e, graph):
    from feram_amd.system import System
    p = batio3_zvr((8, 8, 12), method=method, kelvin=250.0)
    p.bulk_or_film = geometry
    p.epi_strain = -0.005
    p.external_E_field = [0.0, 0.0, 0.001]
    p.E_field_always_zero = False
    p.Q_Nose = 0.1 * p.N / 16 ** 3
    g = System(p)
    g.use_graph(graph)
    rng = np.random.default_rng(3)
    g.put("dipoR", 0.06 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    g.put("dipoP", 0.02 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
    if delta_e:
        g.set_heat_source(delta_e)
    g.md_prologue()
    g.avg_reset(histograms=True)
    g.avg_auto(True)
    return g


@pytest.mark.parametrize("method,geometry,delta_e", CASES)
def test_graph_replay_equals_direct_launches(method, geometry, delta_e):
    print(f"[gpu] step graph {method} {geometry} delta_e={delta_e}", flush=True)
    a = _system(method, geometry, delta_e, True)
    b = _system(method, geometry, delta_e, False)
    for nsteps, kelvin in ((7, None), (5, 320.0), (1, None), (6, None), (22, None)):
        if kelvin is not None:
            a.set_kelvin(kelvin)
            b.set_kelvin(kelvin)
        ea = a.step(nsteps)
        eb = b.step(nsteps)
        assert ea == eb, nsteps
        for f in ("dipoR", "dipoP", "dVddi"):
            assert np.array_equal(a.get(f), b.get(f)), (nsteps, f)
        assert a.nose == b.nose
    sa, sb = a.avg_get(histograms=True), b.avg_get(histograms=True)
    for k in ("u", "uu", "p", "pp", "strain", "n_add"):
        assert np.array_equal(np.array(sa[k]), np.array(sb[k])), k
    assert np.array_equal(sa["distribution"], sb["distribution"])
    if delta_e:
        assert all(np.array_equal(x, y) for x, y in zip(a.layer_energies(True), b.layer_energies(True)))
    a.close()
    b.close()
