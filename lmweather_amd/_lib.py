"""ctypes binding of liblmw_hip.so, the gfx950 library behind include/lmw.h.

Loading fails loudly when the library has not been built: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblmw_hip.so")

LMW_OK, LMW_EINVAL, LMW_ERUNTIME, LMW_EDEVICE = 0, 1, 2, 3
FP64, FP32, BF16X3, BF16X6 = 0, 1, 2, 3
GRID_H, GRID_U, GRID_V, GRID_VORTICITY, GRID_TRACER, GRID_OROGRAPHY = 0, 1, 2, 3, 4, 5
SPEC_ZETA, SPEC_DELTA, SPEC_PHI, SPEC_Q = 0, 1, 2, 3
(DIAG_MEAN_HEIGHT, DIAG_ENERGY, DIAG_KINETIC, DIAG_POTENTIAL, DIAG_ENSTROPHY, DIAG_MAX_WIND, DIAG_ERROR_L2, DIAG_ERROR_LINF,
 DIAG_TIME, DIAG_TRACER_MEAN, DIAG_TRACER_MASS, DIAG_TRACER_MIN, DIAG_TRACER_MAX, DIAG_TRACER_ERROR_L2,
 DIAG_TRACER_ERROR_LINF) = range(15)
POINT_HEIGHT, POINT_U, POINT_V, POINT_VORTICITY, POINT_TRACER = 0, 1, 2, 3, 4
SETUP_WILLIAMSON1, SETUP_WILLIAMSON2, SETUP_WILLIAMSON5, SETUP_WILLIAMSON6 = 1, 2, 5, 6
PLANE_H, PLANE_U, PLANE_V = 0, 1, 2
PDIAG_MEAN_DEPTH, PDIAG_ENERGY, PDIAG_MAX_SPEED, PDIAG_PV_ANOMALY_RMS, PDIAG_TIME = range(5)


class DeviceError(RuntimeError):
    """A HIP failure reported by the library (LMW_EDEVICE)."""


class SphereParams(C.Structure):
    """lmw_sphere_params = reference SphereParams (cpp/sphere.hpp:20-25); defaults are the reference's."""

    _fields_ = [("radius", C.c_double), ("omega", C.c_double), ("gravity", C.c_double), ("diffusion", C.c_double)]

    def __init__(self, radius=6.37122e6, omega=7.292e-5, gravity=9.80616, diffusion=0.0):
        super().__init__(radius, omega, gravity, diffusion)


class PlaneParams(C.Structure):
    """lmw_plane_params = reference PlaneParams (cpp/plane.hpp:9-18); defaults are the reference's."""

    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("lx", C.c_double), ("ly", C.c_double), ("gravity", C.c_double),
                ("depth", C.c_double), ("f0", C.c_double), ("beta", C.c_double), ("viscosity", C.c_double),
                ("drag", C.c_double), ("linear", C.c_int)]

    def __init__(self, nx=0, ny=0, lx=0.0, ly=0.0, gravity=9.81, depth=100.0, f0=0.0, beta=0.0, viscosity=0.0, drag=0.0,
                 linear=False):
        super().__init__(nx, ny, lx, ly, gravity, depth, f0, beta, viscosity, drag, int(bool(linear)))


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)

# name -> argtypes (every function returns lmw_status)
SIGNATURES = {
    "lmw_device_info": [C.c_char_p, C.c_size_t, _ip, C.POINTER(C.c_size_t)],
    "lmw_fft": [_dp, C.c_size_t, C.c_int, C.c_int],
    "lmw_gauss_legendre": [C.c_int, _dp, _dp],
    "lmw_sphere_create": [C.c_int, C.POINTER(SphereParams), C.POINTER(_vp)],
    "lmw_sphere_destroy": [_vp],
    "lmw_sphere_dims": [_vp, _ip, _ip, _ip],
    "lmw_sphere_set_precision": [_vp, C.c_int],
    "lmw_sphere_set_state_grid": [_vp, _dp, _dp, _dp, C.c_size_t],
    "lmw_sphere_set_orography_grid": [_vp, _dp, C.c_size_t],
    "lmw_sphere_set_axis_tilt": [_vp, C.c_double],
    "lmw_sphere_set_diffusion": [_vp, C.c_double],
    "lmw_sphere_mark_reference": [_vp],
    "lmw_sphere_add_mode": [_vp, C.c_int, C.c_int, C.c_double, C.c_double],
    "lmw_sphere_perturb": [_vp, C.c_double, C.c_uint64, C.c_int],
    "lmw_sphere_step": [_vp, C.c_double],
    "lmw_sphere_steps": [_vp, C.c_double, C.c_int64],
    "lmw_sphere_run": [_vp, C.c_double, C.c_double],
    "lmw_sphere_stable_dt": [_vp, C.c_double, _dp],
    "lmw_sphere_get_grid": [_vp, C.c_int, _dp, C.c_size_t],
    "lmw_sphere_get_spectral": [_vp, C.c_int, _dp, C.c_size_t],
    "lmw_sphere_set_spectral": [_vp, C.c_int, _dp, C.c_size_t],
    "lmw_sphere_set_tracer_grid": [_vp, _dp, C.c_size_t],
    "lmw_sphere_set_advect_only": [_vp, C.c_int],
    "lmw_sphere_tracer_flags": [_vp, _ip, _ip],
    "lmw_sphere_diag": [_vp, C.c_int, _dp],
    "lmw_sphere_spectrum": [_vp, _dp, _dp, C.c_size_t],
    "lmw_sphere_point": [_vp, C.c_int, C.c_double, C.c_double, _dp],
    "lmw_sphere_mode_coefficient": [_vp, C.c_int, C.c_int, C.c_int, _dp],
    "lmw_sphere_wave_crest": [_vp, C.c_int, C.c_double, _dp],
    "lmw_sphere_synthesize": [_vp, _dp, _dp],
    "lmw_sphere_analyze": [_vp, _dp, _dp],
    "lmw_sphere_analyze_divergence": [_vp, _dp, _dp, _dp],
    "lmw_sphere_winds": [_vp, _dp, _dp, _dp, _dp],
    "lmw_sphere_tendencies": [_vp, _dp, _dp, _dp],
    "lmw_sphere_gauss": [_vp, _dp, _dp],
    "lmw_sphere_get_tables": [_vp, _dp, _dp],
    "lmw_ensemble_create": [C.c_int, _ip, C.POINTER(SphereParams), C.c_int, C.POINTER(_vp)],
    "lmw_ensemble_destroy": [_vp],
    "lmw_ensemble_member_dims": [_vp, C.c_int, _ip, _ip, _ip],
    "lmw_ensemble_set_state_grid": [_vp, C.c_int, _dp, _dp, _dp, C.c_size_t],
    "lmw_ensemble_set_orography_grid": [_vp, C.c_int, _dp, C.c_size_t],
    "lmw_ensemble_set_axis_tilt": [_vp, C.c_int, C.c_double],
    "lmw_ensemble_set_spectral": [_vp, C.c_int, C.c_int, _dp, C.c_size_t],
    "lmw_ensemble_get_spectral": [_vp, C.c_int, C.c_int, _dp, C.c_size_t],
    "lmw_ensemble_get_grid": [_vp, C.c_int, C.c_int, _dp, C.c_size_t],
    "lmw_ensemble_diag": [_vp, C.c_int, C.c_int, _dp],
    "lmw_ensemble_groups": [_vp, _ip, _ip, _ip],
    "lmw_ensemble_steps": [_vp, _dp, _i64p],
    "lmw_ensemble_steps_graph": [_vp, _dp, _i64p],
    "lmw_ensemble_steps_timed": [_vp, _dp, _i64p, _dp],
    "lmw_ensemble_macro_steps": [_vp, _dp, _ip, C.c_int64],
    "lmw_ensemble_set_persistent": [_vp, C.c_int],
    "lmw_ensemble_autotune": [_vp, C.c_int, _ip, _dp],
    "lmw_sphere_autotune": [_vp, C.c_int, _ip, _dp],
    "lmw_tile_variants": [C.c_int, C.c_int, _ip, _ip],
    "lmw_ensemble_set_tiles": [_vp, C.c_int, _ip],
    "lmw_corrector_create": [C.c_int, C.c_int, C.c_uint64, C.POINTER(_vp)],
    "lmw_corrector_destroy": [_vp],
    "lmw_corrector_dims": [_vp, _ip, _ip, C.POINTER(C.c_size_t)],
    "lmw_corrector_get_params": [_vp, _fp, C.c_size_t],
    "lmw_corrector_set_params": [_vp, _fp, C.c_size_t],
    "lmw_corrector_set_scales": [_vp, _fp, _fp, C.c_size_t],
    "lmw_corrector_loss_grad": [_vp, _fp, _fp, C.c_int, _dp, _fp],
    "lmw_corrector_forward": [_vp, _fp, C.c_int, _fp, _dp],
    "lmw_corrector_forward_noalloc": [_vp, _vp, C.c_int, _vp, _vp],  # device pointers
    "lmw_corrector_set_dataset": [_vp, _fp, _fp, C.c_int],
    "lmw_corrector_train": [_vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _dp],
    "lmw_sphere_corrected_step": [_vp, _vp, C.c_double, C.c_int, C.c_int, C.c_double, _ip, _dp],
    "lmw_ensemble_corrected_step": [_vp, _vp, C.c_double, C.c_int, C.c_int, C.c_double, _ip, _dp],
    "lmw_ensemble_get_truncated": [_vp, C.c_int, _fp, C.c_size_t],
    "lmw_ensemble_set_truncated": [_vp, C.c_int, _fp, C.c_size_t],
    "lmw_plane_create": [C.POINTER(PlaneParams), C.c_int, C.c_int, C.POINTER(_vp)],
    "lmw_plane_destroy": [_vp],
    "lmw_plane_set_params": [_vp, C.POINTER(PlaneParams)],
    "lmw_plane_reset": [_vp],
    "lmw_plane_perturb": [_vp, C.c_int, C.c_double, C.c_uint64],
    "lmw_plane_set_field": [_vp, C.c_int, C.c_int, _dp, C.c_size_t],
    "lmw_plane_get_field": [_vp, C.c_int, C.c_int, _dp, C.c_size_t],
    "lmw_plane_step": [_vp, C.c_double],
    "lmw_plane_steps": [_vp, C.c_double, C.c_int64],
    "lmw_plane_steps_timed": [_vp, C.c_double, C.c_int64, _dp],
    "lmw_plane_run": [_vp, C.c_double, C.c_double],
    "lmw_plane_stable_dt": [_vp, C.c_double, _dp],
    "lmw_plane_diag": [_vp, C.c_int, C.c_int, _dp],
    "lmw_plane_tendency": [_vp, C.c_int, _dp, _dp, _dp],
    "lmw_plane_set_tile": [_vp, C.c_int],
    "lmw_plane_get_params": [_vp, C.POINTER(PlaneParams), _ip, _ip],
    "lmw_plane_zonal_amplitude": [_vp, C.c_int, C.c_int, C.c_int, _dp],
    "lmw_plane_zonal_crest": [_vp, C.c_int, C.c_int, C.c_int, _dp],
    "lmw_plane_v_zonal_crest": [_vp, C.c_int, C.c_int, C.c_int, _dp],
    "lmw_plane_harmonic_energy": [_vp, C.c_int, C.c_int, _dp],
    "lmw_sphere_setup": [_vp, C.c_int, C.c_double],
    "lmw_ensemble_setup": [_vp, C.c_int, C.c_int, C.c_int, C.c_double],
    "lmw_ensemble_set_state_grids": [_vp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_size_t],
    "lmw_sphere_save_csv": [_vp, C.c_char_p],
    "lmw_sphere_save_spectrum": [_vp, C.c_char_p],
    "lmw_sphere_save_zonal_mean": [_vp, C.c_char_p],
    "lmw_plane_save_csv": [_vp, C.c_int, C.c_char_p],
    "lmw_plane_save_row": [_vp, C.c_int, C.c_int, C.c_char_p],
    "lmw_plane_save_column": [_vp, C.c_int, C.c_int, C.c_char_p],
    "lmw_sphere_fields": [_vp, C.c_char_p, C.c_char_p, C.c_char_p],
    "lmw_sphere_orography": [_vp, C.c_char_p],
    "lmw_sphere_tracer": [_vp, C.c_char_p],
    "lmw_ensemble_fields": [_vp, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p],
    "lmw_ensemble_orography": [_vp, C.c_int, C.c_int, C.c_char_p],
    "lmw_plane_fields": [_vp, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p],
}


def text(s: str) -> bytes:
    """An expression string as the bytes the library compiles (UTF-8, NUL-terminated by ctypes)."""
    return s.encode("utf-8") if isinstance(s, str) else bytes(s)

_lib = None


def lib():
    """The loaded library (built by ``python -m lmweather_amd.build``); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m lmweather_amd.build` (no CPU fallback exists)")
        so = C.CDLL(LIB_PATH)
        so.lmw_last_error.restype = C.c_char_p
        so.lmw_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(so, name, None)
            if fn is None:
                continue
            fn.restype = C.c_int
            fn.argtypes = args
        _lib = so
    return _lib


def check(rc: int) -> None:
    """Raise the Python counterpart of the reference's exception (ValueError for std::invalid_argument, RuntimeError
    for std::runtime_error) carrying the library's message."""
    if rc == LMW_OK:
        return
    msg = lib().lmw_last_error().decode()
    if rc == LMW_EINVAL:
        raise ValueError(msg)
    if rc == LMW_EDEVICE:
        raise DeviceError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def as_f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def as_c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)
