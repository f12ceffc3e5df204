import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from feram_amd.params import AMU_IN_EV_A2_PS2, Param  # noqa: E402

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP product)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def load_golden(name="reference_goldens.json"):
    with open(os.path.join(GOLDEN_DIR, name)) as f:
        return json.load(f)


def param_from_dict(d, **override) -> Param:
    """Build a Param from a goldens' parameter dict (same tag names as .feram)."""
    p = Param()
    for k, v in d.items():
        if k == "L":
            p.Lx, p.Ly, p.Lz = v
        elif k == "mass_amu":
            p.mass_dipo = [v / AMU_IN_EV_A2_PS2] * 3
        elif k == "acoustic_mass_amu":
            p.mass_acou = [v / AMU_IN_EV_A2_PS2] * 3
        elif k in ("P_kappa2", "P_alpha", "P_gamma", "P_k1", "P_k2", "P_k3", "P_k4", "Z_star"):
            p.__dict__[k] = [float(v)] * 3
        else:
            setattr(p, k, v)
    for k, v in override.items():
        setattr(p, k, v)
    return p


@pytest.fixture(scope="session")
def goldens():
    return load_golden()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle
