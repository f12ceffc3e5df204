"""fhip_force_static, the allocation-free force evaluation, against fhip_force.

The entry point keeps every intermediate in the library's fixed device
scratch array and works on caller-owned device arrays (allocated here with the
hipMalloc of the HIP runtime the library loaded).  Two calls in a row, the second on a smaller supercell, must each
give fhip_force's dV/du, energies and strain on a fresh context with the same
parameters, u and L(k) table: forces normwise 1e-12, energy terms 1e-12 of
the sum of |terms| (the FFT twiddles are computed on the device instead of
the host, so agreement is to rounding).  A supercell too large for the
scratch array is rejected before any work.
"""
import ctypes as C

import numpy as np
import pytest

from feram_amd.params import batio3_zvr

pytestmark = pytest.mark.gpu

TOL = 1e-12
ENERGY_FIELDS = ["acou_kinetic", "dipo_kinetic", "short_range", "long_range", "dipole_E_field", "unharmonic",
                 "homo_strain", "homo_coupling", "inho_strain", "inho_coupling", "inho_modulation", "total",
                 "H_Nose_Poincare"]


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    loaded (the caller-owned device arrays of the entry point)."""

    H2D, D2H = 1, 2

    def __init__(self):
        from feram_amd import _lib
        _lib.load()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        ptr = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(ptr), max(a.nbytes, 8)) == 0
        self.ptrs.append(ptr)
        assert self.rt.hipMemcpy(ptr, a.ctypes.data, a.nbytes, self.H2D) == 0
        return ptr

    def get(self, ptr, n):
        out = np.empty(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for ptr in self.ptrs:
            self.rt.hipFree(ptr)
        self.ptrs = []


def _static_force(p, lr, u):
    from feram_amd import _lib
    from feram_amd.system import params_struct
    L = _lib.load()
    s = params_struct(p)
    N = p.N
    hip = _Hip()
    try:
        u_d = hip.put(np.moveaxis(u, -1, 0).ravel())
        lr_d = hip.put(np.ravel(lr))
        F_d = hip.put(np.full(3 * N, np.nan))
        e_d = hip.put(np.full(13, np.nan))
        st_d = hip.put(np.full(6, np.nan))
        err = C.create_string_buffer(256)
        rc = L.fhip_force_static(C.byref(s), lr_d, u_d, F_d, e_d, st_d, None, err, 256)
        assert rc == 0, err.value.decode()
        F = np.moveaxis(hip.get(F_d, 3 * N).reshape(3, p.Lz, p.Ly, p.Lx), 0, -1)
        e = dict(zip(ENERGY_FIELDS, hip.get(e_d, 13).tolist()))
        st = hip.get(st_d, 6)
    finally:
        hip.free()
    return F, e, st


def _check(p, seed):
    from feram_amd.system import System
    rng = np.random.default_rng(seed)
    u = 0.08 + 0.05 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3))
    with System(p) as g:
        g.put("dipoR", u)
        e_ref = g.mirror_force_mirror()
        F_ref = g.get("dVddi")
        st_ref = g.strain
        lr = g.get_long_range()
    F, e, st = _static_force(p, lr, u)
    ferr = np.max(np.abs(F - F_ref)) / np.max(np.abs(F_ref))
    scale = sum(abs(e_ref[k]) for k in ENERGY_FIELDS[2:11])
    eerr = max(abs(e[k] - e_ref[k]) for k in ENERGY_FIELDS) / scale
    print(f"[gpu] static force L=({p.Lx},{p.Ly},{p.Lz}): force err {ferr:.2e}, energy err {eerr:.2e}", flush=True)
    assert ferr <= TOL and eerr <= TOL
    assert np.allclose(st, st_ref, rtol=1e-12, atol=1e-16)


def test_static_force_matches_context_force_twice():
    p1 = batio3_zvr((16, 12, 24))
    p1.GPa = [-4.0, -5.0, -6.0]
    p1.external_E_field = [0.0005, 0.0, 0.001]
    p1.E_field_always_zero = False
    _check(p1, 3)
    p2 = batio3_zvr((8, 8, 8))          # the second call on a smaller supercell
    _check(p2, 4)


def test_static_force_rejects_what_does_not_fit():
    from feram_amd import _lib
    from feram_amd.system import params_struct
    L = _lib.load()
    err = C.create_string_buffer(256)
    big = params_struct(batio3_zvr(64))
    dummy = C.c_void_p(1)
    rc = L.fhip_force_static(C.byref(big), dummy, dummy, dummy, dummy, None, None, err, 256)
    assert rc == -2                                  # FHIP_ERR_UNSUPPORTED
    assert b"bytes of scratch" in err.value
    fp32 = params_struct(batio3_zvr(8), precision=1)
    rc = L.fhip_force_static(C.byref(fp32), dummy, dummy, dummy, dummy, None, None, err, 256)
    assert rc == -2 and b"FP64" in err.value
