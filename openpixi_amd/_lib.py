"""ctypes binding of the gfx950 library ``libopx.so`` (C ABI in ``include/opx.h``).

There is no fallback: if the library is missing or cannot be loaded, every use
raises ``OpxLibraryError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopx.so")

OPX_OK = 0
OPX_E_ARG = -1
OPX_E_UNSUPPORTED = -2
OPX_E_DEVICE = -3
OPX_E_STATE = -4
OPX_E_NUMERIC = -5

OPX_FP64 = 0
OPX_FP32 = 1

FIELD_U, FIELD_UNEXT, FIELD_E, FIELD_J, FIELD_RHO = range(5)

OPT_KERNEL, OPT_TILE, OPT_CHUNKS, OPT_GLOBAL_POINTERS, OPT_REUNIT_EVERY = range(5)
KERNEL_AUTO, KERNEL_LINK, KERNEL_STREAM, KERNEL_PLAQ, KERNEL_TILE, KERNEL_PLANE, KERNEL_CELL = range(7)
SOLVER_FAST, SOLVER_IMPLICIT = 0, 1
CGC_CHARGE, CGC_GAUSS = 0, 1
BATCH_OPT_GRID, BATCH_OPT_FORM = range(2)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/opx.h
SIGNATURES = {
    "opx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.POINTER(_vp)]),
    "opx_destroy": (None, [_vp]),
    "opx_last_error": (ctypes.c_char_p, [_vp]),
    "opx_get_info": (ctypes.c_int, [_vp, _i64p, _ip, _ip]),
    "opx_set_stream": (ctypes.c_int, [_vp, _vp]),
    "opx_sync": (ctypes.c_int, [_vp]),
    "opx_upload_field": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "opx_download_field": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "opx_set_links_from_coefficients": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "opx_set_efield_from_coefficients": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    "opx_set_active_region": (ctypes.c_int, [_vp, _ip, _ip]),
    "opx_set_evaluation_region": (ctypes.c_int, [_vp, _ip, _ip]),
    "opx_step": (ctypes.c_int, [_vp, ctypes.c_double]),
    "opx_step_links": (ctypes.c_int, [_vp, ctypes.c_double]),
    "opx_set_field_solver": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "opx_group_pow": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _dp]),
    "opx_store_fields": (ctypes.c_int, [_vp]),
    "opx_reset_current": (ctypes.c_int, [_vp]),
    "opx_reset_charge": (ctypes.c_int, [_vp]),
    "opx_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "opx_get_option": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "opx_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "opx_reunitarize": (ctypes.c_int, [_vp]),
    "opx_gauss_project": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp]),
    "opx_leapfrog": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "opx_measure": (ctypes.c_int, [_vp, _dp]),
    "opx_projected_energy_density": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "opx_projected_energy_density_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "opx_unitarity_test": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]),
    "opx_gauss_restore": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp, _dp]),
    "opx_projected_poynting": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "opx_superparticles_upload": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp,
                                                 ctypes.c_int64]),
    "opx_superparticles_download_charges": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    "opx_interpolate_to_grid": (ctypes.c_int, [_vp, ctypes.c_int]),
    "opx_interpolate_chargedensity": (ctypes.c_int, [_vp, ctypes.c_int]),
    "opx_superparticles_count": (ctypes.c_int, [_vp, _ip, _i64p]),
    "opx_superparticles_download_descriptors": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, _ip]),
    "opx_cgc_mv_charge": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int64, _dp, ctypes.c_int64]),
    "opx_cgc_mv_charge_seeded": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                                _dp, ctypes.c_int64]),
    "opx_java_gaussians": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _dp]),
    "opx_cgc_set_charge": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
    "opx_cgc_lightcone_solve": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "opx_cgc_create_superparticles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip]),
    "opx_cgc_download": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int64]),
    "opx_cgc_wilson_tadpole": (ctypes.c_int, [_vp, _dp]),
    "opx_cgc_wilson_dipole": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _i64p]),
    "opx_cgc_create_particles": (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _ip, _ip]),
    "opx_particles_count": (ctypes.c_int, [_vp, _i64p]),
    "opx_cgc_clear": (ctypes.c_int, [_vp]),
    "opx_simulation_initialize": (ctypes.c_int, [_vp]),
    "opx_simulation_step": (ctypes.c_int, [_vp, ctypes.c_int]),
    "opx_simulation_steps_done": (ctypes.c_int, [_vp]),
    "opx_check_health": (ctypes.c_int, [_vp, ctypes.c_double, _i64p, _i64p]),
    "opx_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(_vp)]),
    "opx_batch_destroy": (None, [_vp]),
    "opx_batch_last_error": (ctypes.c_char_p, [_vp]),
    "opx_batch_event": (_vp, [_vp, ctypes.c_int]),
    "opx_batch_set_stream": (ctypes.c_int, [_vp, _vp]),
    "opx_batch_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "opx_batch_get_option": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "opx_batch_get_info": (ctypes.c_int, [_vp, _ip, _i64p, _ip, _ip]),
    "opx_batch_step_links": (ctypes.c_int, [_vp, ctypes.c_double]),
    "opx_batch_leapfrog": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "opx_batch_measure": (ctypes.c_int, [_vp, _dp]),
    "opx_batch_sync": (ctypes.c_int, [_vp]),
    "opx_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "opx_particles_upload": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int64]),
    "opx_particles_download": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, ctypes.c_int64]),
    "opx_particles_interpolate_chargedensity": (ctypes.c_int, [_vp]),
    "opx_particles_push": (ctypes.c_int, [_vp, ctypes.c_double]),
    "opx_particles_interpolate_to_grid": (ctypes.c_int, [_vp]),
    "opx_particles_reassign": (ctypes.c_int, [_vp]),
    "opx_adjoint_gradient": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                            ctypes.c_int64]),
}


class OpxLibraryError(RuntimeError):
    """The gfx950 library is missing or failed to load."""


class OpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"opx error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libopx.so (raises OpxLibraryError when it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OpxLibraryError(f"{LIB_PATH} is not built; run `python -m openpixi_amd.build`")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - depends on the machine
        raise OpxLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(_ip)


def ivec(vals):
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def dvec(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])
