"""ctypes binding of the C ABI in include/tdse.h (libtdse.so, built in-tree).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the propagator.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libtdse.so")

TDSE_OK = 0
TDSE_E_INVALID = -1
TDSE_E_UNSUPPORTED = -2
TDSE_E_OOM = -3
TDSE_E_DEVICE = -4
TDSE_E_TIMEOUT = -5
TDSE_E_NO_DEVICE = -6
TDSE_E_STATE = -7

TDSE_FP32 = 0
TDSE_FP64 = 1

TDSE_FLAG_UNFUSED_BASELINE = 1
TDSE_FLAG_IMAGINARY_TIME = 2
TDSE_FLAG_LATTICE = 4

TDSE_OBS_NORM = 1
TDSE_OBS_POSITION = 2
TDSE_OBS_ENERGY = 4
TDSE_OBS_MOMENTUM = 8
TDSE_OBS_VARIANCE = 16
TDSE_OBS_PEAK = 32

# every symbol include/tdse.h declares (checked by the CPU test tier)
EXPORTS = (
    "tdse_version", "tdse_status_string", "tdse_last_error",
    "tdse_ctx_create", "tdse_ctx_destroy", "tdse_ctx_bytes_in_use", "tdse_plan_bytes",
    "tdse_plan_create", "tdse_plan_destroy", "tdse_plan_set_potential", "tdse_plan_set_mask",
    "tdse_state_upload", "tdse_state_download", "tdse_state_copy", "tdse_state_device_ptr",
    "tdse_step", "tdse_plan_sync", "tdse_plan_stream",
    "tdse_observe", "tdse_probability_in_slab", "tdse_state_scale",
    "tdse_profile_step", "tdse_plan_sweep_name", "tdse_plan_step_bytes",
    "tdse_plan_autotune", "tdse_plan_get_variants", "tdse_plan_set_variants",
    "tdse_build_harmonic", "tdse_build_barrier", "tdse_build_regularized_coulomb",
    "tdse_build_coulomb_floor", "tdse_build_absorbing_mask", "tdse_build_gaussian",
    "tdse_batch_create", "tdse_batch_destroy", "tdse_batch_step", "tdse_batch_sync",
    "tdse_batch_stream", "tdse_batch_launches_per_round", "tdse_plan_build",
    "tdse_relax_deflated", "tdse_plan_relax_stats", "tdse_build_soft_coulomb",
    "tdse_plan_set_drive", "tdse_plan_set_time", "tdse_plan_get_time", "tdse_plan_set_magnetic",
    "tdse_lattice_set_uniform_field", "tdse_lattice_set_solenoid", "tdse_observe_noalloc",
    "tdse_build_phase_tables", "tdse_state_copy_kernel", "tdse_batch_submit", "tdse_sync",
    "tdse_probability_in_box", "tdse_state_download_f32",
)

TDSE_POT = {"none": 0, "harmonic": 1, "regularized_coulomb": 2, "coulomb_floor": 3, "barrier": 4,
            "soft_coulomb": 5}


class Recipe(C.Structure):
    _fields_ = [
        ("potential", C.c_int),
        ("strength", C.c_double),
        ("center", C.c_double * 3),
        ("x_lo", C.c_double),
        ("x_hi", C.c_double),
        ("mask_width", C.c_double),
        ("r0", C.c_double * 3),
        ("sigma", C.c_double * 3),
        ("k0", C.c_double * 3),
        ("softening", C.c_double),
    ]


class RelaxStats(C.Structure):
    _fields_ = [("norm_sq", C.c_double), ("energy", C.c_double), ("peak", C.c_double)]


class TdseError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


class Observables(C.Structure):
    _fields_ = [
        ("norm", C.c_double),
        ("sum_abs2", C.c_double),
        ("position", C.c_double * 3),
        ("potential", C.c_double),
        ("kinetic", C.c_double),
        ("energy", C.c_double),
        ("sigma", C.c_double * 3),
        ("momentum", C.c_double * 3),
        ("energy_variance", C.c_double),
        ("peak", C.c_double),
    ]


_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib() -> C.CDLL:
    """The loaded libtdse.so (raises if it was never built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python shoeridingerequationsolver_amd/build.py` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        L.tdse_version.restype = C.c_char_p
        L.tdse_status_string.restype = C.c_char_p
        L.tdse_last_error.restype = C.c_char_p
        L.tdse_plan_sweep_name.restype = C.c_char_p
        L.tdse_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]
        L.tdse_ctx_destroy.argtypes = [C.c_void_p]
        L.tdse_ctx_bytes_in_use.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.tdse_plan_bytes.argtypes = [_ip, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
        L.tdse_plan_create.argtypes = [C.c_void_p, _ip, _dp, _dp, C.c_double, C.c_double,
                                       C.c_int, C.c_uint, C.POINTER(C.c_void_p)]
        L.tdse_plan_destroy.argtypes = [C.c_void_p]
        L.tdse_plan_set_potential.argtypes = [C.c_void_p, _dp]
        L.tdse_plan_set_mask.argtypes = [C.c_void_p, _dp]
        L.tdse_state_upload.argtypes = [C.c_void_p, _dp]
        L.tdse_state_download.argtypes = [C.c_void_p, _dp]
        L.tdse_state_copy.argtypes = [C.c_void_p, C.c_void_p]
        L.tdse_state_copy_kernel.argtypes = [C.c_void_p, C.c_void_p]
        L.tdse_state_device_ptr.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.tdse_step.argtypes = [C.c_void_p, C.c_int]
        L.tdse_plan_sync.argtypes = [C.c_void_p, C.c_int]
        L.tdse_plan_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.tdse_observe.argtypes = [C.c_void_p, C.c_uint, C.POINTER(Observables)]
        L.tdse_probability_in_slab.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.tdse_probability_in_box.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.tdse_state_download_f32.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.tdse_state_scale.argtypes = [C.c_void_p, C.c_double]
        L.tdse_profile_step.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _ip, _ip]
        L.tdse_plan_sweep_name.argtypes = [C.c_void_p, C.c_int]
        L.tdse_plan_step_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.tdse_plan_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.tdse_plan_get_variants.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]
        L.tdse_plan_set_variants.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
        for name in ("tdse_build_harmonic", "tdse_build_regularized_coulomb",
                     "tdse_build_coulomb_floor"):
            getattr(L, name).argtypes = [_ip, _dp, _dp, C.c_double, _dp, _dp]
        L.tdse_build_barrier.argtypes = [_ip, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.tdse_build_absorbing_mask.argtypes = [_ip, _dp, _dp, C.c_double, _dp]
        L.tdse_build_gaussian.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.tdse_batch_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int,
                                        C.POINTER(C.c_void_p)]
        L.tdse_batch_destroy.argtypes = [C.c_void_p]
        L.tdse_batch_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.tdse_batch_sync.argtypes = [C.c_void_p, C.c_int]
        L.tdse_batch_submit.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                        C.c_int]
        L.tdse_sync.argtypes = [C.c_void_p, C.c_int]
        L.tdse_batch_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.tdse_batch_launches_per_round.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.tdse_plan_build.argtypes = [C.c_void_p, C.POINTER(Recipe)]
        L.tdse_relax_deflated.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int]
        L.tdse_plan_relax_stats.argtypes = [C.c_void_p, C.POINTER(RelaxStats)]
        L.tdse_build_soft_coulomb.argtypes = [_ip, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.tdse_plan_set_drive.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double]
        L.tdse_plan_set_time.argtypes = [C.c_void_p, C.c_double]
        L.tdse_plan_get_time.argtypes = [C.c_void_p, _dp]
        L.tdse_plan_set_magnetic.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.tdse_lattice_set_uniform_field.argtypes = [C.c_void_p, C.c_double]
        L.tdse_lattice_set_solenoid.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                                C.c_int]
        L.tdse_observe_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _ip, _dp, _dp,
                                           C.c_void_p, C.c_void_p]
        L.tdse_build_phase_tables.argtypes = [_ip, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        _LIB = L
    return _LIB


def check(status: int) -> None:
    if status != TDSE_OK:
        L = lib()
        msg = L.tdse_last_error().decode() or L.tdse_status_string(status).decode()
        raise TdseError(status, msg)


def ivec(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def dvec(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def dptr(a):
    return a.ctypes.data_as(_dp)
