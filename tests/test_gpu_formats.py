"""write_system / .restart round trip of device state (feram_amd.formats).

After a few NP steps the state is written in the reference's .coord format
(p divided by s_Nose, the optimised w reconstructed as coord_module.F:457-469
does) and read into a fresh context the way initialize_dipoR reads a
.restart file; u and p must come back to the printed precision, and the
restarted run's first force evaluation must match the original's.
"""
import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu


def test_coord_restart_round_trip(tmp_path):
    from feram_amd import formats
    from feram_amd.system import System
    print("[gpu] .coord / .restart round trip 8^3", flush=True)
    p = batio3_zvr(8, method="md", kelvin=300.0)
    g = System(p)
    rng = np.random.default_rng(3)
    g.put("dipoR", 0.08 + 0.05 * rng.standard_normal((8, 8, 8, 3)))
    g.md_prologue()
    g.nose_poincare(5)
    path = tmp_path / "run.0000000005.coord"
    formats.write_coord(g, str(path))
    lines = path.read_text().splitlines()
    assert len(lines) == 512 and all(len(l) == 14 + 30 + 9 * 14 + 39 for l in lines)
    s = g.nose[0]
    h = System(p)
    fields = formats.read_restart(h, str(path))
    assert np.max(np.abs(h.get("dipoR") - g.get("dipoR"))) <= 5e-7
    pd = g.get("dipoP") / s
    assert np.all(np.abs(h.get("dipoP") - pd) <= 5e-6 * np.abs(pd) + 1e-300)
    assert np.allclose(fields["acouR"], g.get("acouR_optimised"), rtol=5e-6, atol=1e-300)
    h.md_prologue()
    g2 = System(p)
    g2.put("dipoR", fields["dipoR"])
    eg2 = g2.mirror_force_mirror()
    assert np.allclose(h.get("dVddi"), g2.get("dVddi"), rtol=0, atol=1e-12)
    for x in (g, h, g2):
        x.close()
