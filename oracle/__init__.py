"""ctypes front end of the CPU FP64 oracle (liboracle.so, built from feram_oracle.c).

TEST INFRASTRUCTURE ONLY.  This package is the checker the HIP product is
compared with.  Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product package ``feram_amd`` never
does.  Parity pins for every oracle function are listed in the tests that
exercise them (tests/test_oracle_*.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

METHODS = {"md": 0, "vs": 1, "lf": 2}
GEOMETRIES = {"bulk": 0, "film": 1, "epit": 2, "e110": 3}


class FoParams(C.Structure):
    _fields_ = [
        ("Lx", C.c_int), ("Ly", C.c_int), ("Lz", C.c_int),
        ("method", C.c_int), ("bulk_or_film", C.c_int), ("gap_id", C.c_int),
        ("E_field_always_zero", C.c_int), ("pad_", C.c_int),
        ("GPa", C.c_double * 3), ("kelvin", C.c_double),
        ("mass_dipo", C.c_double * 3), ("mass_acou", C.c_double * 3),
        ("Q_Nose", C.c_double), ("a0", C.c_double), ("epi_strain", C.c_double),
        ("dt", C.c_double), ("external_E_field", C.c_double * 3),
        ("P_kappa2", C.c_double * 3), ("P_alpha", C.c_double * 3), ("P_gamma", C.c_double * 3),
        ("P_k1", C.c_double * 3), ("P_k2", C.c_double * 3), ("P_k3", C.c_double * 3),
        ("P_k4", C.c_double * 3), ("j", C.c_double * 7),
        ("B11", C.c_double), ("B12", C.c_double), ("B44", C.c_double),
        ("B1xx", C.c_double), ("B1yy", C.c_double), ("B4yz", C.c_double),
        ("modulation_constant", C.c_double), ("Z_star", C.c_double * 3),
        ("epsilon_inf", C.c_double), ("kappa", C.c_double), ("gap_dipole_u", C.c_double * 3),
    ]


ENERGY_FIELDS = ["acou_kinetic", "dipo_kinetic", "short_range", "long_range", "dipole_E_field",
                 "unharmonic", "homo_strain", "homo_coupling", "inho_strain", "inho_coupling",
                 "inho_modulation", "total", "H_Nose_Poincare"]


class FoEnergy(C.Structure):
    _fields_ = [(n, C.c_double) for n in ENERGY_FIELDS]

    def as_dict(self):
        return {n: getattr(self, n) for n in ENERGY_FIELDS}


class FoAverage(C.Structure):
    """fo_average: Average_type sums (average_module.F:17-30)."""
    _fields_ = [("energy", FoEnergy), ("e2", C.c_double), ("acou_kinetic_true", C.c_double),
                ("dipo_kinetic_true", C.c_double), ("strain", C.c_double * 6), ("u", C.c_double * 3),
                ("uu", C.c_double * 6), ("p", C.c_double * 3), ("pp", C.c_double * 6),
                ("u_tmp", C.c_double * 3), ("uu_tmp", C.c_double * 3), ("p_tmp", C.c_double * 3),
                ("pp_tmp", C.c_double * 3), ("n_add", C.c_int64)]

    def as_dict(self) -> dict:
        d = {"energy": self.energy.as_dict()}
        for name, _ in self._fields_[1:]:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "feram_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.fo_create.restype = vp
        L.fo_create.argtypes = [C.POINTER(FoParams)]
        L.fo_destroy.argtypes = [vp]
        L.fo_setup_tables.argtypes = [vp, C.c_void_p]
        L.fo_setup_tables.restype = C.c_int
        for n in ("fo_kappa", "fo_target_kinetic_energy", "fo_temperature_eV"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [vp]
        L.fo_pNa03.restype = C.c_double
        L.fo_pNa03.argtypes = [vp, C.c_int]
        L.fo_field.restype = C.POINTER(C.c_double)
        L.fo_field.argtypes = [vp, C.c_char_p]
        for n in ("fo_strain", "fo_homogeneous", "fo_nose"):
            getattr(L, n).restype = C.POINTER(C.c_double)
            getattr(L, n).argtypes = [vp]
        L.fo_relaxor.restype = C.POINTER(C.c_int)
        L.fo_relaxor.argtypes = [vp]
        L.fo_sum_p.restype = C.POINTER(C.c_ubyte)
        L.fo_sum_p.argtypes = [vp]
        L.fo_recount_sum_p.argtypes = [vp]
        for n in ("fo_mirror_force_mirror", "fo_nose_poincare", "fo_velocity_scaling",
                  "fo_nose_poincare_hamiltonian", "fo_md_prologue"):
            getattr(L, n).argtypes = [vp, C.POINTER(FoEnergy)]
        L.fo_leapfrog.argtypes = [vp, C.POINTER(FoEnergy), C.c_int]
        for n in ("fo_kinetic_energy_dipo", "fo_kinetic_energy_acou"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [vp]
        L.fo_sum_potential_energies.restype = C.c_double
        L.fo_add_to_average.argtypes = [vp, C.POINTER(FoEnergy), C.POINTER(FoAverage), vp, vp]
        L.fo_add_to_average.restype = None
        L.fo_sum_potential_energies.argtypes = [C.POINTER(FoEnergy)]
        L.fo_init_normal.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                     C.c_int, C.POINTER(C.c_double)]
        L.fo_uni64_seed.restype = C.c_double
        L.fo_uni64_seed.argtypes = [C.c_int, C.c_int]
        L.fo_uni64.restype = C.c_double
        L.fo_uni64_fill.argtypes = [C.c_void_p, C.c_long]
        L.fo_normal_dist_seed.restype = C.c_double
        L.fo_normal_dist_seed.argtypes = [C.c_int, C.c_int]
        L.fo_normal_dist.restype = C.c_double
        L.fo_normal_dist.argtypes = [C.c_double, C.c_double]
        for n in ("fo_r2c", "fo_c2r"):
            getattr(L, n).argtypes = [vp, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.fo_fft1d.argtypes = [C.c_void_p, C.c_int, C.c_int]
        for n in ("fo_B_decay", "fo_C_decay"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_double, C.c_double, C.c_double]
        L.fo_wave.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_long, C.c_long, C.c_long, C.c_long,
                              C.POINTER(C.c_double)]
        L.fo_set_heat_source.argtypes = [vp, C.c_double, C.c_double]
        L.fo_set_heat_source.restype = None
        L.fo_layers.argtypes = [vp, C.c_int]
        L.fo_layers.restype = C.POINTER(C.c_double)
        L.fo_wave.restype = None
        _lib = L
    return _lib


def params_struct(p) -> FoParams:
    """Fill the oracle's parameter struct from a ``feram_amd.params.Param``-like object."""
    s = FoParams()
    s.Lx, s.Ly, s.Lz = p.Lx, p.Ly, p.Lz
    s.method = METHODS[p.method]
    s.bulk_or_film = GEOMETRIES[p.bulk_or_film]
    s.gap_id = p.gap_id
    s.E_field_always_zero = 1 if p.E_field_always_zero else 0
    for name in ("GPa", "mass_dipo", "mass_acou", "external_E_field", "P_kappa2", "P_alpha",
                 "P_gamma", "P_k1", "P_k2", "P_k3", "P_k4", "j", "Z_star", "gap_dipole_u"):
        arr = getattr(s, name)
        for i, v in enumerate(getattr(p, name)):
            arr[i] = float(v)
    for name in ("kelvin", "Q_Nose", "a0", "epi_strain", "dt", "B11", "B12", "B44", "B1xx",
                 "B1yy", "B4yz", "modulation_constant", "epsilon_inf", "kappa"):
        setattr(s, name, float(getattr(p, name)))
    return s


class Oracle:
    """One oracle context: the reference's ``Coord_type`` plus its tables."""

    REAL3 = ("dipoR", "dipoP", "dVddi", "acouR", "acouP", "dVdacR", "E_field", "dVddi_long", "Z_u_R")

    def __init__(self, p, modulation_R=None, tables=True):
        self.p = p
        self._s = params_struct(p)
        self._L = lib()
        self.ctx = self._L.fo_create(C.byref(self._s))
        self.N = p.Lx * p.Ly * p.Lz
        self.Kx = p.Lx // 2 + 1
        self.Kh = self.Kx * p.Ly * p.Lz
        self._mod = None
        if tables:
            self.setup_tables(modulation_R)

    def setup_tables(self, modulation_R=None):
        ptr = None
        if modulation_R is not None:
            self._mod = np.ascontiguousarray(modulation_R, dtype=np.float64).reshape(-1)
            ptr = self._mod.ctypes.data
        rc = self._L.fo_setup_tables(self.ctx, ptr)
        if rc != 0:
            raise ValueError(f"elastic(): modulation error (stop {rc})")

    def close(self):
        if self.ctx:
            self._L.fo_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- array views ----------------------------------------------------
    def _view(self, name, count):
        ptr = self._L.fo_field(self.ctx, name.encode())
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(count,))

    def field(self, name):
        p = self.p
        shape3 = (p.Lz, p.Ly, p.Lx)
        if name in self.REAL3:
            return self._view(name, 3 * self.N).reshape(shape3 + (3,))
        if name in ("didiR", "inhoR"):
            return self._view(name, 6 * self.N).reshape(shape3 + (6,))
        if name in ("Z_starR", "mass_dipo_inv", "mass_acou_inv"):
            return self._view(name, self.N).reshape(shape3)
        kshape = (p.Lz, p.Ly, self.Kx)
        if name in ("long_range", "short_range", "inhomo_K"):
            return self._view(name, 9 * self.Kh).reshape(kshape + (3, 3))   # [.., b, a]
        if name == "inhomo_K_G":
            return self._view(name, 12 * self.Kh).reshape(kshape + (4, 3))  # [.., b, a]
        if name == "inhomo_coup2_K":
            return self._view(name, 18 * self.Kh).reshape(kshape + (6, 3))  # [.., i, a]
        ncomp = {"acouK": 3, "dipoK": 3, "didiK": 6, "c_k_modulation_k": 3, "modulation_k": 1}[name]
        v = self._view(name, 2 * ncomp * self.Kh)
        return v.view(np.complex128).reshape(kshape + (ncomp,))

    @property
    def relaxor(self):
        return np.ctypeslib.as_array(self._L.fo_relaxor(self.ctx), shape=(self.N,)).reshape(
            self.p.Lz, self.p.Ly, self.p.Lx)

    @property
    def sum_p(self):
        return np.ctypeslib.as_array(self._L.fo_sum_p(self.ctx), shape=(self.N,)).reshape(
            self.p.Lz, self.p.Ly, self.p.Lx)

    def recount_sum_p(self):
        self._L.fo_recount_sum_p(self.ctx)

    def set_heat_source(self, delta_e, delta_e_acou=0.0):
        """Heat-source / sink layers of leapfrog (leapfrog.F:38-83)."""
        self._L.fo_set_heat_source(self.ctx, float(delta_e), float(delta_e_acou))

    def layers(self):
        """(e_layer_dipo, e_layer_acou) of the last leapfrog step, Lz values each."""
        return tuple(np.ctypeslib.as_array(self._L.fo_layers(self.ctx, w), shape=(self.p.Lz,)).copy()
                     for w in (0, 1))

    @property
    def strain(self):
        return np.ctypeslib.as_array(self._L.fo_strain(self.ctx), shape=(6,))

    @property
    def nose(self):
        """View of (s_Nose, pi_Nose, H0)."""
        return np.ctypeslib.as_array(self._L.fo_nose(self.ctx), shape=(3,))

    @property
    def kappa(self):
        return self._L.fo_kappa(self.ctx)

    # ---- hot path -------------------------------------------------------
    def mirror_force_mirror(self, e=None):
        e = e if e is not None else FoEnergy()
        self._L.fo_mirror_force_mirror(self.ctx, C.byref(e))
        return e

    def md_prologue(self):
        e = FoEnergy()
        self._L.fo_md_prologue(self.ctx, C.byref(e))
        return e

    def nose_poincare(self, e):
        self._L.fo_nose_poincare(self.ctx, C.byref(e))
        return e

    def leapfrog(self, e, i_step=1):
        self._L.fo_leapfrog(self.ctx, C.byref(e), int(i_step))
        return e

    def velocity_scaling(self, e):
        self._L.fo_velocity_scaling(self.ctx, C.byref(e))
        return e

    def step(self, e):
        m = self.p.method
        if m == "md":
            return self.nose_poincare(e)
        if m == "lf":
            return self.leapfrog(e, 1)
        return self.velocity_scaling(e)

    def add_to_average(self, e, acc, dist1=None, dist3=None):
        """add_to_Average (average_module.F:57-163) into acc (FoAverage); dist1
        int64 (199, 3) [k + 99][component], dist3 int32 (199, 199, 199) [kz][ky][kx]."""
        self._L.fo_add_to_average(self.ctx, C.byref(e), C.byref(acc),
                                  dist1.ctypes.data_as(C.c_void_p) if dist1 is not None else None,
                                  dist3.ctypes.data_as(C.c_void_p) if dist3 is not None else None)
        return acc

    def kinetic_energy_dipo(self):
        return self._L.fo_kinetic_energy_dipo(self.ctx)

    def kinetic_energy_acou(self):
        return self._L.fo_kinetic_energy_acou(self.ctx)

    def init_normal(self, avg, dev, seed, k_stripe=(0.0, 0.0, 0.0)):
        a = (C.c_double * 3)(*avg)
        d = (C.c_double * 3)(*dev)
        k = (C.c_double * 3)(*k_stripe)
        self._L.fo_init_normal(self.ctx, a, d, int(seed[0]), int(seed[1]), k)

    # ---- FFT helpers (tests) -------------------------------------------
    def r2c(self, real):
        real = np.ascontiguousarray(real, dtype=np.float64)
        out = np.zeros(2 * self.Kh)
        self._L.fo_r2c(self.ctx, real.ctypes.data, 1, 0, out.ctypes.data)
        return out.view(np.complex128).reshape(self.p.Lz, self.p.Ly, self.Kx)

    def c2r(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        out = np.zeros(self.N)
        self._L.fo_c2r(self.ctx, spec.ctypes.data, 1, 0, out.ctypes.data)
        return out.reshape(self.p.Lz, self.p.Ly, self.p.Lx)


def fft1d(x, sign=-1):
    x = np.array(x, dtype=np.complex128)
    lib().fo_fft1d(x.ctypes.data, len(x), int(sign))
    return x


def uni64_seed(s1, s2):
    return lib().fo_uni64_seed(int(s1), int(s2))


def uni64():
    return lib().fo_uni64()


def uni64_fill(n):
    out = np.empty(int(n))
    lib().fo_uni64_fill(out.ctypes.data, int(n))
    return out


def normal_dist_seed(s1, s2):
    return lib().fo_normal_dist_seed(int(s1), int(s2))


def normal_dist(mu, sigma):
    return lib().fo_normal_dist(float(mu), float(sigma))


WAVES = {"triangular_cos": 1, "triangular_sin": 2, "ramping_on": 3, "ramping_off": 4, "hysteresis": 5}


def wave(kind: str, amplitude, period: int, n_thermalize: int, n_average: int, i_step: int) -> np.ndarray:
    """fo_wave: the uniform field of step i_step (wave_module.F:8-67, hysteresis_loop.F:47-49)."""
    a = (C.c_double * 3)(*[float(v) for v in amplitude])
    out = (C.c_double * 3)()
    lib().fo_wave(WAVES[kind], a, int(period), int(n_thermalize), int(n_average), int(i_step), out)
    return np.array(out[:])
