"""Statistical parity with the reference's own BaTiO3 phase-transition scan.

The reference ships the averaged output of its 16^3 heating scan
(tests/golden/bto16_phase_transition.json, copied from
03example-BaTiO3-bulk-phase-transition/16x16x16-40000-10000/heating.avg):
rhombohedral up to ~215 K, orthorhombic to ~255 K, tetragonal to ~315 K,
cubic above.  The same chained protocol (feram_amd.sweeps.temperature_chain:
each temperature starts from the previous final state with p / s_Nose, as the
reference's .coord -> .restart hand-over does) runs here for four seeds in one
replica batch, shortened to 10000 + 3000 steps per temperature and 10 K
spacing.  Deep inside each phase at least three of the four chains must be in
the reference's phase (a shortened protocol can leave one superheated), and
over those chains the sorted |<u>| of the polar components and the sorted
normal strains must agree with the reference's single run within three times
the combined error: the standard error of our mean and the seed-to-seed
deviation that one reference run carries (floors 0.004 A and 3e-4).
"""
import json
import os

import numpy as np
import pytest

from feram_amd.sweeps import phase_of

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEEP = {170: "R", 180: "R", 190: "R", 200: "R", 240: "O",
        270: "T", 280: "T", 290: "T", 300: "T", 340: "C", 350: "C"}


def test_heating_scan_matches_reference_example():
    from feram_amd.sweeps import temperature_chain
    ref = json.load(open(os.path.join(HERE, "golden", "bto16_phase_transition.json")))["heating"]
    ref = {r[0]: r for r in ref}
    temps = np.arange(170, 351, 10)
    print(f"[gpu] heating chain 16^3, 4 seeds, {len(temps)} temperatures x 13000 steps", flush=True)

    def progress(i, T, u):
        print(f"[gpu]   {T:.0f} K  " + "  ".join(
            "/".join(f"{x:.3f}" for x in np.sort(np.abs(v))) for v in u), flush=True)

    res = temperature_chain(temps, seeds=(1, 2, 3, 4), n_thermalize=10000, n_average=3000, progress=progress)
    strain = np.sort(res["strain"][:, :, :3], axis=2)
    for i, T in enumerate(res["temps"]):
        if T not in DEEP:
            continue
        r = ref[T]
        u_ref = np.sort(np.abs(r[7:10]))
        e_ref = np.sort(r[1:4])
        assert phase_of(u_ref) == DEEP[T]
        ok = [j for j, ph in enumerate(res["phases"][i]) if ph == DEEP[T]]
        assert len(ok) >= 3, (T, res["phases"][i])
        us = res["u_sorted"][i][ok]
        m = us.mean(axis=0)
        sd = us.std(axis=0, ddof=1)
        polar = u_ref > 0.04
        tol = np.maximum(3 * sd * np.sqrt(1 + 1 / len(ok)), 0.004)
        assert np.all(np.abs(m - u_ref)[polar] <= tol[polar]), (T, m, u_ref)
        st = strain[i][ok]
        sm, ssd = st.mean(axis=0), st.std(axis=0, ddof=1)
        assert np.all(np.abs(sm - e_ref) <= np.maximum(3 * ssd * np.sqrt(1 + 1 / len(ok)), 3e-4)), (T, sm, e_ref)
