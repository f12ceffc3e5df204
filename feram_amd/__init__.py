"""MI355X-native per-step force-and-integrate path of feram.

The step (dipole-dipole k-space convolution through in-house LDS FFTs,
on-site and strain forces, acoustic path, Nose-Poincare / leapfrog /
velocity-scaling update) runs as HIP kernels for gfx950 behind the C ABI in
include/feram_hip.h.  ``System`` is the Python host side of that ABI;
``Param`` mirrors feram's Param_type and reads .feram inputs.
"""
from .params import Param, batio3_zvr  # noqa: F401


def __getattr__(name):
    # System loads the HIP library lazily, so that the parameter layer imports
    # on machines without the built library.
    if name in ("System", "Batch", "FhipError"):
        from . import system
        return getattr(system, name)
    raise AttributeError(name)
