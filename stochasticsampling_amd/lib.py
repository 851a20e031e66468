"""Loader and ctypes signatures of libss.so (include/ss_api.h).

The library is built in-tree (``python -m stochasticsampling_amd.build``) and
loaded from ``stochasticsampling_amd/_lib/libss.so``.  There is no fallback:
if the library is missing, every entry point raises.
"""
import ctypes as C
import os

from .params import ERRORS, SS_OK, Box, Info, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libss.so")

_P = C.POINTER
_d = _P(C.c_double)
_lib = None


class SSError(RuntimeError):
    """A negative return code of libss, with its ss_last_error() message."""

    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _declare(L):
    pp = _P(Params)
    vp = C.c_void_p
    sig = {
        "ss_params_default": (None, [pp]),
        "ss_last_error": (C.c_char_p, []),
        "ss_version": (C.c_int, []),
        "ss_device_count": (C.c_int, [_P(C.c_int)]),
        "ss_create": (C.c_int, [pp, _P(vp)]),
        "ss_destroy": (C.c_int, [vp]),
        "ss_init": (C.c_int, [vp, _d, C.c_uint64]),
        "ss_set_noise": (C.c_int, [vp, _d, C.c_uint64]),
        "ss_step": (C.c_int, [vp, C.c_uint64, _P(C.c_uint64)]),
        "ss_get_timestep": (C.c_int, [vp, _P(C.c_uint64)]),
        "ss_get_particles": (C.c_int, [vp, _d, C.c_uint64]),
        "ss_get_distribution": (C.c_int, [vp, _P(C.c_uint32), _d]),
        "ss_get_flow_field": (C.c_int, [vp, _d, _d]),
        "ss_get_magnetic_field": (C.c_int, [vp, _d, _d]),
        "ss_get_snapshot": (C.c_int, [vp, _d, _d, _P(C.c_uint64)]),
        "ss_resume": (C.c_int, [vp, _d, _d, C.c_uint64, C.c_uint64]),
        "ss_get_rng_state": (C.c_int, [vp, _P(C.c_uint64), C.c_int32]),
        "ss_set_rng_state": (C.c_int, [vp, _P(C.c_uint64), C.c_int32]),
        "ss_get_noise": (C.c_int, [vp, _d, C.c_uint64]),
        "ss_profile_steps": (C.c_int, [vp, C.c_uint64, _d, C.c_char_p, C.c_int, C.c_int]),
        "ss_synchronize": (C.c_int, [vp]),
        "ss_resort": (C.c_int, [vp]),
        "ss_prepare_steps": (C.c_int, [vp, C.c_uint64]),
        "ss_get_diagnostics": (C.c_int, [vp, _d]),
        "ss_get_info": (C.c_int, [vp, _P(Info)]),
        "ss_ensemble_create": (C.c_int, [pp, C.c_int, _P(vp)]),
        "ss_ensemble_destroy": (C.c_int, [vp]),
        "ss_ensemble_member": (C.c_int, [vp, C.c_int, _P(vp)]),
        "ss_ensemble_step": (C.c_int, [vp, C.c_uint64, _P(C.c_uint64)]),
        "ss_ensemble_profile_steps": (C.c_int, [vp, C.c_uint64, _d, C.c_char_p, C.c_int, C.c_int]),
        "ss_initial_condition": (C.c_int, [pp, C.c_double, C.c_uint64, _d, C.c_uint64]),
        "ss_create_isotropic": (C.c_int, [C.c_uint64, _P(Box), C.c_uint64, C.c_int32, _d]),
        "ss_create_homogeneous": (C.c_int, [C.c_uint64, C.c_double, _P(Box), C.c_uint64, C.c_int32, _d]),
        "ss_fft_plan_create": (
            C.c_int,
            [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P(vp)],
        ),
        "ss_fft_execute": (C.c_int, [vp, vp, vp]),
        "ss_fft_execute_host": (C.c_int, [vp, vp]),
        "ss_fft_execute_host_noalloc": (C.c_int, [vp, vp]),
        "ss_fft_plan_destroy": (C.c_int, [vp]),
        "ss_flow_solve": (C.c_int, [pp, _d, _d, _d]),
        "ss_solver_create": (C.c_int, [pp, C.c_int, _P(vp)]),
        "ss_solver_destroy": (C.c_int, [vp]),
        "ss_solver_mean_flow_field": (C.c_int, [vp, C.c_double, _d, _d, _d]),
        "ss_solver_mean_magnetic_field": (C.c_int, [vp, _d, _d, _d]),
        "ss_keep_distribution": (C.c_int, [vp, C.c_int]),
        "ss_magnetic_solve": (C.c_int, [pp, _d, _d, _d]),
        "ss_density_gradient": (C.c_int, [pp, _d, _d]),
        "ss_interpolate_vector_field": (C.c_int, [pp, _d, _d, C.c_uint64, _d]),
        "ss_debug_update": (C.c_int, [pp, C.c_uint64, _d, _d, _d, _d]),
        "ss_debug_coord_to_grid": (C.c_int, [pp, C.c_uint64, _d, _P(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return sig


EXPORTS = None


def load():
    """Load libss.so once; raises if it has not been built."""
    global _lib, EXPORTS
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m stochasticsampling_amd.build`"
            )
        L = C.CDLL(LIB_PATH)
        EXPORTS = sorted(_declare(L))
        _lib = L
    return _lib


def check(rc):
    """Raise SSError for a negative return code."""
    if rc < SS_OK:
        msg = load().ss_last_error().decode(errors="replace")
        raise SSError(rc, msg)
    return rc


def device_count():
    n = C.c_int(0)
    check(load().ss_device_count(C.byref(n)))
    return n.value


def dptr(a):
    return a.ctypes.data_as(_d)
