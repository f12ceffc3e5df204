"""Host-side mirror of the reference's per-step interface on the HIP library.

``System`` plays the role of feram's ``Coord_type`` plus its step routines:
its methods are the subroutines the reference's drivers call
(``mirror_force_mirror``, ``nose_poincare``, ``leapfrog``,
``velocity_scaling``; see include/feram_hip.h for the file:line each one
replaces) and return the ``Energy_type`` values as a dict.  State lives on the
GPU between calls; ``get``/``put`` move fields in the reference's component
order as numpy arrays of shape (Lz, Ly, Lx, 3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .params import GEOMETRIES, METHODS, Param


class FhipError(RuntimeError):
    pass


def params_struct(p: Param, precision: int = 0) -> _lib.FhipParams:
    s = _lib.FhipParams()
    s.Lx, s.Ly, s.Lz, s.padding_y = p.Lx, p.Ly, p.Lz, p.padding_y
    s.method = METHODS[p.method]
    s.bulk_or_film = GEOMETRIES[p.bulk_or_film]
    s.precision = precision
    s.E_field_always_zero = 1 if p.E_field_always_zero else 0
    s.gap_id = int(p.gap_id)
    for name in ("dt", "kelvin", "Q_Nose", "a0", "epi_strain", "B11", "B12", "B44", "B1xx",
                 "B1yy", "B4yz", "modulation_constant", "epsilon_inf", "kappa"):
        setattr(s, name, float(getattr(p, name)))
    for name in ("GPa", "mass_dipo", "mass_acou", "P_kappa2", "P_alpha", "P_gamma", "P_k1",
                 "P_k2", "P_k3", "P_k4", "j", "Z_star", "external_E_field"):
        arr = getattr(s, name)
        for i, v in enumerate(getattr(p, name)):
            arr[i] = float(v)
    return s


class System:
    """One device context (``fhip_ctx``) for a supercell."""

    def __init__(self, p: Param, *, precision: int = 0):
        self.p = p
        self._L = _lib.load()
        self._s = params_struct(p, precision)
        ctx = C.c_void_p()
        rc = self._L.fhip_create(C.byref(self._s), C.byref(ctx))
        self.ctx = ctx
        if rc != 0:
            msg = self._L.fhip_last_error(ctx).decode() if ctx.value else ""
            self.close()
            raise FhipError(f"fhip_create: {_lib.ERRORS.get(rc, rc)}: {msg}")
        self.N = p.Lx * p.Ly * p.Lz

    # ---- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self._L.fhip_destroy(self.ctx)
        self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc < 0:
            raise FhipError(f"{what}: {_lib.ERRORS.get(rc, rc)}: {self._L.fhip_last_error(self.ctx).decode()}")
        return rc

    # ---- fields ------------------------------------------------------------
    def put(self, name: str, arr) -> None:
        p = self.p
        a = np.asarray(arr, dtype=np.float64).reshape(p.Lz, p.Ly, p.Lx, 3)
        soa = np.ascontiguousarray(np.moveaxis(a, -1, 0))
        self._check(self._L.fhip_put_soa(self.ctx, _lib.FIELDS[name], soa.ctypes.data), f"put {name}")

    def get(self, name: str) -> np.ndarray:
        p = self.p
        soa = np.empty((3, p.Lz, p.Ly, p.Lx))
        self._check(self._L.fhip_get_soa(self.ctx, _lib.FIELDS[name], soa.ctypes.data), f"get {name}")
        return np.ascontiguousarray(np.moveaxis(soa, 0, -1))

    def put_padded(self, name: str, host: np.ndarray) -> None:
        """Host array in the reference's padded layout (coord_module.F:134-141)."""
        h = np.ascontiguousarray(host, dtype=np.float64)
        self._check(self._L.fhip_put(self.ctx, _lib.FIELDS[name], h.ctypes.data), f"put {name}")

    def get_padded(self, name: str, host: np.ndarray) -> np.ndarray:
        assert host.flags.c_contiguous and host.dtype == np.float64
        self._check(self._L.fhip_get(self.ctx, _lib.FIELDS[name], host.ctypes.data), f"get {name}")
        return host

    def padded_shape(self, ncomp: int = 3):
        """Shape of a padded host field in C order (z, y+pad, 2(Lx/2+1), ncomp)."""
        p = self.p
        return (p.Lz, p.Ly + p.padding_y, 2 * (p.Lx // 2 + 1), ncomp)

    def set_long_range(self, table: np.ndarray) -> None:
        """c%long_range(3,3,0:Lx/2,0:Ly+pad-1,0:Lz-1) given in C order (z, y+pad, kx, b, a)."""
        t = np.ascontiguousarray(table, dtype=np.float64)
        self._check(self._L.fhip_set_long_range(self.ctx, t.ctypes.data), "set_long_range")

    def get_long_range(self) -> np.ndarray:
        """The device table as c%long_range in C order (z, y+pad, kx, b, a)."""
        p = self.p
        t = np.zeros((p.Lz, p.Ly + p.padding_y, p.Lx // 2 + 1, 3, 3))
        self._check(self._L.fhip_get_long_range(self.ctx, t.ctypes.data), "get_long_range")
        return t

    def set_modulation(self, modulation_R: np.ndarray) -> None:
        """modulation_R of shape (Lz, Ly, Lx) (elastic.F:187-200)."""
        p = self.p
        h = np.zeros((p.Lz, p.Ly + p.padding_y, 2 * (p.Lx // 2 + 1)))
        h[:, :p.Ly, :p.Lx] = modulation_R
        self._check(self._L.fhip_set_modulation(self.ctx, h.ctypes.data), "set_modulation")

    def set_cells(self, relaxor=None, sum_p=None, mass_dipo_inv=None) -> None:
        """Per-cell species (1..3), sum_p mask and 1/m, each of shape (Lz, Ly, Lx)."""
        p = self.p
        shp = (p.Lz, p.Ly + p.padding_y, 2 * (p.Lx // 2 + 1))
        keep = []

        def pad(a, dtype):
            if a is None:
                return None
            h = np.zeros(shp, dtype=dtype)
            h[:, :p.Ly, :p.Lx] = a
            keep.append(h)
            return h.ctypes.data

        rc = self._L.fhip_set_cells(self.ctx, pad(relaxor, np.int32), pad(sum_p, np.int32),
                                    pad(mass_dipo_inv, np.float64))
        self._check(rc, "set_cells")
        self._sum_p = None if sum_p is None else np.asarray(sum_p).astype(bool)

    @property
    def film(self) -> bool:
        return self.p.bulk_or_film in ("film", "epit")

    def counted(self) -> np.ndarray:
        """The sum_p mask the device uses, (Lz, Ly, Lx): the one given to
        set_cells (all cells by default) with, for film/epit, the upper half
        of z and the dead layers removed (read_defects_u.F:52-69)."""
        p = self.p
        m = getattr(self, "_sum_p", None)
        m = np.ones((p.Lz, p.Ly, p.Lx), bool) if m is None else m.copy()
        if self.film:
            m[(p.Lz + 1) // 2:] = False
            if p.gap_id >= 1:
                m[0] = False
            if p.gap_id == 2:
                m[p.Lz // 2] = False
        return m

    def mean_u(self) -> np.ndarray:
        """<u> over the sum_p cells (hysteresis_loop.F:85-89)."""
        return self.get("dipoR")[self.counted()].mean(axis=0)

    def avg_u_norm(self) -> int:
        """Divisor of the u, u^2 sums per averaging step: n_sum_p for film/epit,
        N otherwise (average_module.F:206-215)."""
        return int(self.counted().sum()) if self.film else self.N

    # ---- scalars -----------------------------------------------------------
    @property
    def nose(self):
        s, pi, h0 = C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.fhip_get_nose(self.ctx, C.byref(s), C.byref(pi), C.byref(h0)), "get_nose")
        return s.value, pi.value, h0.value

    def set_nose(self, s, pi, H0):
        self._check(self._L.fhip_set_nose(self.ctx, float(s), float(pi), float(H0)), "set_nose")

    @property
    def strain(self) -> np.ndarray:
        out = (C.c_double * 6)()
        self._check(self._L.fhip_get_strain(self.ctx, out), "get_strain")
        return np.array(out[:])

    def set_E_field_add(self, E) -> None:
        e = (C.c_double * 3)(*[float(v) for v in E])
        self._check(self._L.fhip_set_E_field_add(self.ctx, e), "set_E_field_add")

    E_SCHEDULES = {"constant": 0, "triangular_cos": 1, "triangular_sin": 2, "ramping_on": 3,
                   "ramping_off": 4, "hysteresis": 5}

    def set_E_schedule(self, kind: str, amplitude, period: int = 0, n_thermalize: int = 0,
                       n_average: int = 0, i_step: int = 1) -> None:
        """Device-side E_add(i_step): an E_wave_type of wave_module.F, the
        hysteresis-loop sweep ('hysteresis') or 'constant'."""
        a = (C.c_double * 3)(*[float(v) for v in amplitude])
        self._check(self._L.fhip_set_E_schedule(self.ctx, self.E_SCHEDULES[kind], a, int(period),
                                                int(n_thermalize), int(n_average), int(i_step)), "set_E_schedule")

    def step_info(self):
        """(index of the next step, the E_add it will use)."""
        i = C.c_int64()
        e = (C.c_double * 3)()
        self._check(self._L.fhip_get_step(self.ctx, C.byref(i), e), "get_step")
        return int(i.value), np.array(e[:])

    def set_heat_source(self, delta_e: float, delta_e_acou: float = 0.0) -> None:
        """Heat-source / sink layers of leapfrog (leapfrog.F:38-83)."""
        self._check(self._L.fhip_set_heat_source(self.ctx, float(delta_e), float(delta_e_acou)), "set_heat_source")

    def layer_energies(self, averaged: bool = False):
        """(e_layer_dipo, e_layer_acou) of the last step, or their sums over
        the averaging steps."""
        Lz = self.p.Lz
        d, a = np.zeros(Lz), np.zeros(Lz)
        self._check(self._L.fhip_get_layer_energies(self.ctx, d.ctypes.data, a.ctypes.data, 1 if averaged else 0),
                    "get_layer_energies")
        return d, a

    def set_kelvin(self, kelvin: float) -> None:
        self._check(self._L.fhip_set_kelvin(self.ctx, float(kelvin)), "set_kelvin")
        self.p = self.p.copy(kelvin=float(kelvin))

    # ---- exact resume ---------------------------------------------------------
    def checkpoint(self) -> bytes:
        """The context's state as the library's binary checkpoint
        (fhip_checkpoint_save): loading it into a System made with the same
        parameters resumes bitwise."""
        n = int(self._L.fhip_checkpoint_bytes(self.ctx))
        if n < 0:
            raise FhipError(f"checkpoint_bytes: {_lib.ERRORS.get(n, n)}")
        buf = (C.c_ubyte * n)()
        self._check(self._L.fhip_checkpoint_save(self.ctx, buf, n), "checkpoint_save")
        return bytes(buf)

    def restore(self, data: bytes) -> None:
        """Load a checkpoint written by checkpoint() (fhip_checkpoint_load)."""
        buf = (C.c_ubyte * len(data)).from_buffer_copy(data)
        self._check(self._L.fhip_checkpoint_load(self.ctx, buf, len(data)), "checkpoint_load")

    def save_checkpoint(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    def load_checkpoint(self, path) -> None:
        with open(path, "rb") as f:
            self.restore(f.read())

    # ---- the step ------------------------------------------------------------
    def mirror_force_mirror(self) -> dict:
        e = _lib.FhipEnergy()
        self._check(self._L.fhip_force(self.ctx, C.byref(e)), "force")
        return e.as_dict()

    def md_prologue(self) -> dict:
        e = _lib.FhipEnergy()
        self._check(self._L.fhip_md_prologue(self.ctx, C.byref(e)), "md_prologue")
        return e.as_dict()

    def nose_poincare(self, nsteps: int = 1) -> dict:
        e = _lib.FhipEnergy()
        self._check(self._L.fhip_nose_poincare(self.ctx, int(nsteps), C.byref(e)), "nose_poincare")
        return e.as_dict()

    def leapfrog(self, i_step: int = 1, nsteps: int = 1) -> dict:
        e = _lib.FhipEnergy()
        self._check(self._L.fhip_leapfrog(self.ctx, int(i_step), int(nsteps), C.byref(e)), "leapfrog")
        return e.as_dict()

    def velocity_scaling(self, nsteps: int = 1) -> dict:
        e = _lib.FhipEnergy()
        self._check(self._L.fhip_velocity_scaling(self.ctx, int(nsteps), C.byref(e)), "velocity_scaling")
        return e.as_dict()

    def step(self, nsteps: int = 1) -> dict:
        m = self.p.method
        if m == "md":
            return self.nose_poincare(nsteps)
        if m == "lf":
            return self.leapfrog(1, nsteps)
        return self.velocity_scaling(nsteps)

    # ---- measurement ---------------------------------------------------------
    # ---- averages (add_to_Average, average_module.F:32-163) -----------------
    def avg_reset(self, histograms: bool = False, dipoRavg: bool = False) -> None:
        opts = (1 if histograms else 0) | (2 if dipoRavg else 0)
        self._check(self._L.fhip_avg_reset(self.ctx, opts), "avg_reset")

    def avg_add(self) -> None:
        self._check(self._L.fhip_avg_add(self.ctx), "avg_add")

    def avg_auto(self, enable: bool = True) -> None:
        self._check(self._L.fhip_avg_auto(self.ctx, 1 if enable else 0), "avg_auto")

    def avg_get(self, histograms: bool = False) -> dict:
        """The Average_type sums; with histograms, also distribution (3, 199)
        and distribution3d (199, 199, 199) indexed [kz][ky][kx] + 99."""
        from ._lib import FhipAverage
        a = FhipAverage()
        d1 = d3 = None
        if histograms:
            d1 = np.zeros((199, 3), dtype=np.int64)          # distribution(3, -99:99), component fastest
            d3 = np.zeros((199, 199, 199), dtype=np.int32)   # distribution3d(kx, ky, kz), kx fastest
        self._check(self._L.fhip_avg_get(self.ctx, C.byref(a),
                                         d1.ctypes.data_as(C.c_void_p) if d1 is not None else None,
                                         d3.ctypes.data_as(C.c_void_p) if d3 is not None else None), "avg_get")
        out = a.as_dict()
        if histograms:
            out["distribution"] = d1.T.copy()       # [component][k + 99]
            out["distribution3d"] = d3
        return out

    def profile(self, enable: bool = True) -> None:
        self._check(self._L.fhip_profile(self.ctx, 1 if enable else 0), "profile")

    def use_graph(self, enable: bool = True) -> None:
        """Replay the middle steps of a multi-step call as a captured hipGraph."""
        self._check(self._L.fhip_use_graph(self.ctx, 1 if enable else 0), "use_graph")

    def kernel_stats(self) -> dict:
        n = 32
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        k = self._check(self._L.fhip_kernel_stats(self.ctx, n, names, ms, cnt), "kernel_stats")
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k) if cnt[i] > 0}

    def synchronize(self) -> None:
        self._check(self._L.fhip_synchronize(self.ctx), "synchronize")

    @property
    def device_bytes(self) -> int:
        return int(self._L.fhip_device_bytes(self.ctx))


class Batch:
    """Replicas stepped together (``fhip_batch``): independent ``System``
    objects -- e.g. one per temperature and field of a sweep -- whose steps run
    as one launch per stage and size class, with the middle steps of a call
    replayed from a captured HIP graph.  Each replica evolves exactly as its own
    ``System.step`` would; the systems stay usable between batch calls."""

    def __init__(self, systems):
        self.systems = list(systems)
        self._L = _lib.load()
        arr = (C.c_void_p * len(self.systems))(*[s.ctx.value for s in self.systems])
        b = C.c_void_p()
        rc = self._L.fhip_batch_create(arr, len(self.systems), C.byref(b))
        self.b = b
        if rc != 0:
            msg = self._L.fhip_batch_last_error(b).decode() if b.value else ""
            self.close()
            raise FhipError(f"fhip_batch_create: {_lib.ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            self._L.fhip_batch_destroy(self.b)
        self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc < 0:
            raise FhipError(f"{what}: {_lib.ERRORS.get(rc, rc)}: {self._L.fhip_batch_last_error(self.b).decode()}")
        return rc

    def step(self, nsteps: int = 1) -> list:
        """nsteps steps of every replica; the energies of each after the last."""
        e = (_lib.FhipEnergy * len(self.systems))()
        self._check(self._L.fhip_batch_step(self.b, int(nsteps), e), "batch_step")
        return [x.as_dict() for x in e]

    def use_graph(self, enable: bool = True) -> None:
        self._check(self._L.fhip_batch_graph(self.b, 1 if enable else 0), "batch_graph")

    def tune(self, enable: bool = True) -> None:
        """Time the tile shapes of each size class at its first step (default on)."""
        self._check(self._L.fhip_batch_tune(self.b, 1 if enable else 0), "batch_tune")

    def plan(self) -> list:
        """Per size class: edges, chosen x rows and y lines, measured us per launch."""
        out = []
        for k in range(self.classes):
            shape = (C.c_int32 * 5)()
            us = (C.c_double * 4)()
            tuned = self._L.fhip_batch_plan(self.b, k, shape, us)
            if tuned < 0:
                raise FhipError(f"batch_plan: status {tuned}")
            out.append({"L": tuple(shape[:3]), "x_rows": shape[3], "y_lines": shape[4], "tuned": bool(tuned),
                        "x_us": (us[0], us[1]), "y_us": (us[2], us[3])})
        return out

    @property
    def classes(self) -> int:
        return int(self._L.fhip_batch_classes(self.b))
