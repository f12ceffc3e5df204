"""ctypes declarations of the C ABI in include/feram_hip.h.

The library is loaded from the package directory (built in-tree by
``python -m feram_amd.build``).  There is no fallback: without the HIP
library the package cannot run the step and says so.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libferam_hip.so")

FHIP_OK = 0
ERRORS = {-1: "FHIP_ERR_ARG", -2: "FHIP_ERR_UNSUPPORTED", -3: "FHIP_ERR_HIP",
          -4: "FHIP_ERR_ALLOC", -5: "FHIP_ERR_STATE"}

FIELDS = {"dipoR": 0, "dipoP": 1, "dVddi": 2, "acouR": 3, "acouP": 4, "dVdacR": 5,
          "E_field_local": 6, "acouR_optimised": 7, "dipoRavg": 8}

# Every symbol include/feram_hip.h declares.
EXPORTS = ("fhip_create", "fhip_destroy", "fhip_last_error", "fhip_set_long_range",
           "fhip_get_long_range",
           "fhip_set_modulation", "fhip_set_cells", "fhip_put", "fhip_get", "fhip_put_soa",
           "fhip_get_soa", "fhip_get_nose", "fhip_set_nose", "fhip_get_strain",
           "fhip_set_E_field_add", "fhip_set_kelvin", "fhip_set_heat_source", "fhip_get_layer_energies",
           "fhip_set_E_schedule", "fhip_get_step", "fhip_force", "fhip_force_static", "fhip_md_prologue",
           "fhip_nose_poincare",
           "fhip_leapfrog", "fhip_velocity_scaling", "fhip_profile", "fhip_use_graph", "fhip_kernel_stats",
           "fhip_synchronize", "fhip_copy_rate", "fhip_checkpoint_bytes", "fhip_checkpoint_save",
           "fhip_checkpoint_load", "fhip_device_bytes", "fhip_avg_reset", "fhip_avg_add", "fhip_avg_auto",
           "fhip_avg_get", "fhip_batch_create", "fhip_batch_destroy", "fhip_batch_last_error",
           "fhip_batch_step", "fhip_batch_graph", "fhip_batch_tune", "fhip_batch_plan", "fhip_batch_size",
           "fhip_batch_classes",
           "fhip_mlp_nparams", "fhip_mlp_create", "fhip_mlp_destroy", "fhip_mlp_last_error",
           "fhip_mlp_set_params", "fhip_mlp_get_params", "fhip_mlp_forward", "fhip_mlp_gradient",
           "fhip_mlp_train")


class FhipParams(C.Structure):
    _fields_ = [
        ("Lx", C.c_int32), ("Ly", C.c_int32), ("Lz", C.c_int32), ("padding_y", C.c_int32),
        ("method", C.c_int32), ("bulk_or_film", C.c_int32), ("precision", C.c_int32),
        ("E_field_always_zero", C.c_int32),
        ("dt", C.c_double), ("kelvin", C.c_double), ("Q_Nose", C.c_double), ("a0", C.c_double),
        ("epi_strain", C.c_double),
        ("GPa", C.c_double * 3), ("mass_dipo", C.c_double * 3), ("mass_acou", C.c_double * 3),
        ("P_kappa2", C.c_double * 3), ("P_alpha", C.c_double * 3), ("P_gamma", C.c_double * 3),
        ("P_k1", C.c_double * 3), ("P_k2", C.c_double * 3), ("P_k3", C.c_double * 3),
        ("P_k4", C.c_double * 3), ("j", C.c_double * 7),
        ("B11", C.c_double), ("B12", C.c_double), ("B44", C.c_double),
        ("B1xx", C.c_double), ("B1yy", C.c_double), ("B4yz", C.c_double),
        ("modulation_constant", C.c_double), ("Z_star", C.c_double * 3),
        ("epsilon_inf", C.c_double), ("kappa", C.c_double), ("external_E_field", C.c_double * 3),
        ("gap_id", C.c_int32), ("reserved", C.c_int32),
    ]


ENERGY_FIELDS = ("acou_kinetic", "dipo_kinetic", "short_range", "long_range", "dipole_E_field",
                 "unharmonic", "homo_strain", "homo_coupling", "inho_strain", "inho_coupling",
                 "inho_modulation", "total", "H_Nose_Poincare")


class FhipEnergy(C.Structure):
    _fields_ = [(n, C.c_double) for n in ENERGY_FIELDS]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in ENERGY_FIELDS}


_lib = None


class FhipAverage(C.Structure):
    """fhip_average: Average_type sums (average_module.F:17-30)."""
    _fields_ = [("energy", FhipEnergy), ("e2", C.c_double), ("acou_kinetic_true", C.c_double),
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


def load():
    """Load libferam_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m feram_amd.build` "
                           "(the step has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.fhip_create.argtypes = [C.POINTER(FhipParams), C.POINTER(vp)]
    L.fhip_create.restype = ip
    L.fhip_destroy.argtypes = [vp]
    L.fhip_destroy.restype = None
    L.fhip_last_error.argtypes = [vp]
    L.fhip_last_error.restype = C.c_char_p
    L.fhip_set_long_range.argtypes = [vp, vp]
    L.fhip_get_long_range.argtypes = [vp, vp]
    L.fhip_set_modulation.argtypes = [vp, vp]
    L.fhip_set_cells.argtypes = [vp, vp, vp, vp]
    for n in ("fhip_put", "fhip_get", "fhip_put_soa", "fhip_get_soa"):
        getattr(L, n).argtypes = [vp, ip, vp]
    L.fhip_get_nose.argtypes = [vp, dp, dp, dp]
    L.fhip_set_nose.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.fhip_get_strain.argtypes = [vp, dp]
    L.fhip_set_E_field_add.argtypes = [vp, dp]
    L.fhip_set_kelvin.argtypes = [vp, C.c_double]
    L.fhip_set_kelvin.restype = ip
    L.fhip_set_heat_source.argtypes = [vp, C.c_double, C.c_double]
    L.fhip_set_heat_source.restype = ip
    L.fhip_get_layer_energies.argtypes = [vp, vp, vp, ip]
    L.fhip_get_layer_energies.restype = ip
    L.fhip_set_E_schedule.argtypes = [vp, ip, dp, ip, ip, ip, C.c_int64]
    L.fhip_set_E_schedule.restype = ip
    L.fhip_get_step.argtypes = [vp, C.POINTER(C.c_int64), dp]
    L.fhip_get_step.restype = ip
    L.fhip_force.argtypes = [vp, C.POINTER(FhipEnergy)]
    L.fhip_force_static.argtypes = [C.POINTER(FhipParams), vp, vp, vp, vp, vp, vp, C.c_char_p, ip]
    L.fhip_force_static.restype = ip
    L.fhip_md_prologue.argtypes = [vp, C.POINTER(FhipEnergy)]
    L.fhip_nose_poincare.argtypes = [vp, ip, C.POINTER(FhipEnergy)]
    L.fhip_leapfrog.argtypes = [vp, ip, ip, C.POINTER(FhipEnergy)]
    L.fhip_velocity_scaling.argtypes = [vp, ip, C.POINTER(FhipEnergy)]
    L.fhip_profile.argtypes = [vp, ip]
    L.fhip_use_graph.argtypes = [vp, ip]
    L.fhip_kernel_stats.argtypes = [vp, ip, C.POINTER(C.c_char_p), dp, C.POINTER(C.c_int64)]
    L.fhip_synchronize.argtypes = [vp]
    L.fhip_checkpoint_bytes.argtypes = [vp]
    L.fhip_checkpoint_bytes.restype = C.c_int64
    L.fhip_checkpoint_save.argtypes = [vp, vp, C.c_int64]
    L.fhip_checkpoint_load.argtypes = [vp, vp, C.c_int64]
    L.fhip_copy_rate.argtypes = [C.c_int64, ip, dp, dp]
    L.fhip_copy_rate.restype = ip
    L.fhip_device_bytes.argtypes = [vp]
    L.fhip_device_bytes.restype = C.c_int64
    L.fhip_avg_reset.argtypes = [vp, ip]
    L.fhip_avg_add.argtypes = [vp]
    L.fhip_avg_auto.argtypes = [vp, ip]
    L.fhip_avg_get.argtypes = [vp, C.POINTER(FhipAverage), vp, vp]
    L.fhip_batch_create.argtypes = [C.POINTER(vp), ip, C.POINTER(vp)]
    L.fhip_batch_destroy.argtypes = [vp]
    L.fhip_batch_destroy.restype = None
    L.fhip_batch_last_error.argtypes = [vp]
    L.fhip_batch_last_error.restype = C.c_char_p
    L.fhip_batch_step.argtypes = [vp, ip, C.POINTER(FhipEnergy)]
    L.fhip_batch_graph.argtypes = [vp, ip]
    L.fhip_batch_tune.argtypes = [vp, ip]
    L.fhip_batch_plan.argtypes = [vp, ip, C.POINTER(C.c_int32), dp]
    L.fhip_batch_size.argtypes = [vp]
    L.fhip_batch_classes.argtypes = [vp]
    for n in ("fhip_batch_create", "fhip_batch_step", "fhip_batch_graph", "fhip_batch_tune", "fhip_batch_plan",
              "fhip_batch_size", "fhip_batch_classes"):
        getattr(L, n).restype = ip
    fp = C.POINTER(C.c_float)
    L.fhip_mlp_nparams.argtypes = []
    L.fhip_mlp_create.argtypes = [C.c_uint64, C.POINTER(vp)]
    L.fhip_mlp_destroy.argtypes = [vp]
    L.fhip_mlp_destroy.restype = None
    L.fhip_mlp_last_error.argtypes = [vp]
    L.fhip_mlp_last_error.restype = C.c_char_p
    L.fhip_mlp_set_params.argtypes = [vp, fp]
    L.fhip_mlp_get_params.argtypes = [vp, fp]
    L.fhip_mlp_forward.argtypes = [vp, ip, fp, fp]
    L.fhip_mlp_gradient.argtypes = [vp, ip, fp, fp, fp, fp, fp]
    L.fhip_mlp_train.argtypes = [vp, ip, fp, fp, fp, ip, C.c_float, fp]
    for n in ("fhip_mlp_nparams", "fhip_mlp_create", "fhip_mlp_set_params", "fhip_mlp_get_params",
              "fhip_mlp_forward", "fhip_mlp_gradient", "fhip_mlp_train"):
        getattr(L, n).restype = ip
    for n in ("fhip_set_long_range", "fhip_get_long_range", "fhip_set_modulation", "fhip_set_cells", "fhip_put",
              "fhip_get", "fhip_put_soa", "fhip_get_soa", "fhip_get_nose", "fhip_set_nose",
              "fhip_get_strain", "fhip_set_E_field_add", "fhip_force", "fhip_md_prologue",
              "fhip_nose_poincare", "fhip_leapfrog", "fhip_velocity_scaling", "fhip_profile", "fhip_use_graph",
              "fhip_kernel_stats", "fhip_synchronize", "fhip_avg_reset", "fhip_avg_add", "fhip_avg_auto",
              "fhip_avg_get"):
        getattr(L, n).restype = ip
    _lib = L
    return L
