"""The 48^3 acoustic heating scan: BaTiO3 bulk 48x48x48, FP64 plain mode,
velocity scaling with acoustic MD (acoustic_mass_amu 46.44, as the
reference's 26example-BaTiO3-acoustic-MD inputs), 200 -> 400 K in 10 K
steps, 2000 thermalising + 2000 averaging steps per temperature, each started
from the previous final state (feram_amd.sweeps.temperature_chain), one
seeded chain.

The scan must pass through the reference's sequence of phases for this
parameter set (rhombohedral, orthorhombic, tetragonal, cubic; the reference's
own 16^3 heating.avg places the transitions near 215, 255 and 315 K):
deep inside each phase the chain is in that phase, the ferroelectric
components of <u> have the size the reference's scan gives there (within
0.01 A), and in the cubic phase the three normal strains agree within 1e-4
(cubic symmetry) and <u> vanishes.  A heating scan this short lags the
transitions by 10-20 K, so only temperatures where both scans are deep
inside the same phase are checked.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# temperatures where this scan and the reference's 16^3 heating.avg are deep
# inside the same phase (the reference's data end at 350 K)
DEEP = {200: "R", 210: "R", 250: "O", 290: "T", 300: "T", 350: "C", 360: "C", 400: "C"}


def test_acoustic_heating_scan_48():
    from feram_amd.params import batio3_zvr
    from feram_amd.sweeps import temperature_chain
    ref = json.load(open(os.path.join(HERE, "golden", "bto16_phase_transition.json")))["heating"]
    ref = {r[0]: r for r in ref}
    temps = np.arange(200, 401, 10)
    base = batio3_zvr(48, method="vs", kelvin=200.0, acoustic_mass_amu=46.44)
    print(f"[gpu] 48^3 acoustic vs heating chain, {len(temps)} temperatures x 4000 steps", flush=True)
    res = temperature_chain(temps, seeds=(1,), base=base, n_thermalize=2000, n_average=2000)
    for i, T in enumerate(res["temps"]):
        us = res["u_sorted"][i, 0]
        st = np.sort(res["strain"][i, 0, :3])
        assert np.all(np.isfinite(us)) and np.all(np.isfinite(st)), T
        if T not in DEEP:
            continue
        ph = DEEP[T]
        assert res["phases"][i][0] == ph, (T, res["phases"][i][0], us)
        r = ref.get(T, ref[min(ref, key=lambda t: abs(t - T))])
        ref_us = np.sort(np.abs(r[7:10]))
        npol = "CTOR".index(ph)
        if npol:
            assert np.allclose(us[3 - npol:], ref_us[3 - npol:], atol=0.01), (T, us, ref_us)
        else:
            assert np.all(us < 0.005), (T, us)
            assert st[2] - st[0] < 1e-4, (T, st)
