"""ctypes binding of libfsamd.so (include/fake_spectra_amd.h).

The library is required: every entry point of this package raises if it is
missing or no GPU is visible.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FSAMD_LIB") or os.path.join(HERE, "_lib", "libfsamd.so")

FSAMD_OK = 0
FSAMD_ERR_ARG = 1
FSAMD_ERR_OOM = 2
FSAMD_ERR_HIP = 3
FSAMD_ERR_UNSUPPORTED = 4
MODE_BATCHED = 0
MODE_PLAIN = 1

# Every symbol include/fake_spectra_amd.h declares.
EXPORTS = ("fsamd_last_error", "fsamd_version", "fsamd_plan_create", "fsamd_plan_destroy",
           "fsamd_plan_npairs", "fsamd_plan_pairs", "fsamd_plan_near_lines", "fsamd_tau",
           "fsamd_colden", "fsamd_particle_interpolate", "fsamd_set_stats", "fsamd_last_stats", "fsamd_last_ms",
           "fsamd_voigt_peak", "fsamd_set_group_size", "fsamd_plan_keep_windows", "fsamd_tau_adjoint",
           "fsamd_tau_autotune", "fsamd_plan_overflow", "fsamd_mean_flux_scale", "fsamd_res_corr",
           "fsamd_flux_power", "fsamd_flux_hist", "fsamd_last_kernel_ms", "fsamd_set_block_waves",
           "fsamd_voigt_eval", "fsamd_plan_window_generation", "fsamd_set_fp64",
           "fsamd_particle_interpolate_noalloc", "fsamd_near_lines", "fsamd_plan_assign_cells",
           "fsamd_pool_release")


class FsamdLine(C.Structure):
    _fields_ = [("lambda_cm", C.c_double), ("gamma", C.c_double), ("fosc", C.c_double),
                ("amumass", C.c_double), ("dens", C.c_void_p), ("out", C.c_void_p)]


class FsamdLineGrad(C.Structure):
    _fields_ = [("lambda_cm", C.c_double), ("gamma", C.c_double), ("fosc", C.c_double),
                ("amumass", C.c_double), ("dens", C.c_void_p), ("grad_tau", C.c_void_p),
                ("grad_dens", C.c_void_p)]


_lock = threading.Lock()
_lib = None


def _declare(L):
    vp, d, i, ll = C.c_void_p, C.c_double, C.c_int, C.c_longlong
    L.fsamd_last_error.argtypes = []
    L.fsamd_last_error.restype = C.c_char_p
    L.fsamd_version.argtypes = []
    L.fsamd_version.restype = i
    L.fsamd_plan_create.argtypes = [d, vp, vp, ll, vp, vp, i, vp, C.POINTER(vp)]
    L.fsamd_plan_create.restype = i
    L.fsamd_plan_destroy.argtypes = [vp]
    L.fsamd_plan_destroy.restype = i
    L.fsamd_plan_npairs.argtypes = [vp]
    L.fsamd_plan_npairs.restype = ll
    L.fsamd_plan_pairs.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.fsamd_plan_pairs.restype = i
    L.fsamd_plan_near_lines.argtypes = [vp, vp, C.POINTER(ll), vp]
    L.fsamd_plan_near_lines.restype = i
    L.fsamd_plan_assign_cells.argtypes = [vp, vp, vp]
    L.fsamd_plan_assign_cells.restype = i
    L.fsamd_near_lines.argtypes = [d, vp, vp, ll, vp, vp, i, vp, C.POINTER(ll), vp]
    L.fsamd_near_lines.restype = i
    L.fsamd_tau.argtypes = [vp, i, i, d, d, vp, vp, vp, vp, C.POINTER(FsamdLine), i, i, vp]
    L.fsamd_tau.restype = i
    L.fsamd_colden.argtypes = [vp, i, i, d, vp, vp, C.POINTER(FsamdLine), i, vp]
    L.fsamd_colden.restype = i
    L.fsamd_particle_interpolate.argtypes = [i, i, i, d, d, d, d, d, d, d, d, vp, vp, vp, vp, vp, ll, vp, vp,
                                             i, vp, vp]
    L.fsamd_particle_interpolate.restype = i
    L.fsamd_particle_interpolate_noalloc.argtypes = L.fsamd_particle_interpolate.argtypes
    L.fsamd_particle_interpolate_noalloc.restype = i
    L.fsamd_set_stats.argtypes = [i]
    L.fsamd_set_stats.restype = i
    L.fsamd_last_stats.argtypes = [C.POINTER(ll), i]
    L.fsamd_last_stats.restype = i
    L.fsamd_last_ms.argtypes = []
    L.fsamd_last_ms.restype = d
    L.fsamd_last_kernel_ms.argtypes = []
    L.fsamd_last_kernel_ms.restype = d
    L.fsamd_voigt_peak.argtypes = [d, d, C.POINTER(d), vp]
    L.fsamd_voigt_peak.restype = i
    L.fsamd_voigt_eval.argtypes = [vp, ll, d, vp, vp]
    L.fsamd_voigt_eval.restype = i
    L.fsamd_set_group_size.argtypes = [i]
    L.fsamd_set_group_size.restype = i
    L.fsamd_set_block_waves.argtypes = [i]
    L.fsamd_set_block_waves.restype = i
    L.fsamd_plan_keep_windows.argtypes = [vp, i]
    L.fsamd_plan_keep_windows.restype = i
    L.fsamd_plan_window_generation.argtypes = [vp, C.POINTER(ll)]
    L.fsamd_plan_window_generation.restype = i
    L.fsamd_set_fp64.argtypes = [i]
    L.fsamd_set_fp64.restype = i
    L.fsamd_pool_release.argtypes = []
    L.fsamd_pool_release.restype = i
    L.fsamd_tau_adjoint.argtypes = [vp, i, i, d, vp, vp, vp, vp, C.POINTER(FsamdLineGrad), i, vp, vp, vp]
    L.fsamd_tau_adjoint.restype = i
    L.fsamd_tau_autotune.argtypes = [vp, i, i, d, d, vp, vp, vp, vp, C.POINTER(FsamdLine), i, vp, C.POINTER(i),
                                     C.POINTER(i), C.POINTER(C.c_float)]
    L.fsamd_tau_autotune.restype = i
    L.fsamd_plan_overflow.argtypes = [vp, C.POINTER(i), vp]
    L.fsamd_plan_overflow.restype = i
    L.fsamd_mean_flux_scale.argtypes = [vp, ll, d, d, d, C.POINTER(d), vp]
    L.fsamd_mean_flux_scale.restype = i
    L.fsamd_res_corr.argtypes = [vp, vp, i, i, d, vp]
    L.fsamd_res_corr.restype = i
    L.fsamd_flux_power.argtypes = [vp, i, i, d, d, d, vp, vp]
    L.fsamd_flux_power.restype = i
    L.fsamd_flux_hist.argtypes = [vp, ll, d, i, vp, vp]
    L.fsamd_flux_hist.restype = i


def lib():
    """Load libfsamd.so once; raises RuntimeError if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError("libfsamd.so is not built (%s); run python -m fake_spectra_amd.build" % LIB_PATH)
            L = C.CDLL(LIB_PATH)
            _declare(L)
            _lib = L
    return _lib


def set_group_size(g: int):
    """Pixels per lane group of the tau kernel: 16, 32 or 64."""
    check(lib().fsamd_set_group_size(int(g)))


def set_block_waves(w: int):
    """Wavefronts per (sightline, line) work item of the tau kernel: 2, 4 or 8."""
    check(lib().fsamd_set_block_waves(int(w)))


def check(rc: int):
    """Map a libfsamd status to the reference's exception classes."""
    if rc == FSAMD_OK:
        return
    msg = lib().fsamd_last_error().decode(errors="replace")
    if rc == FSAMD_ERR_ARG or rc == FSAMD_ERR_UNSUPPORTED:
        raise ValueError(msg)
    if rc == FSAMD_ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError("libfsamd error %d: %s" % (rc, msg))
