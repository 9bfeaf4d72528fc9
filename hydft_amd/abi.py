"""ctypes binding of libhydft (include/hydft.h).

The library is the product; this module only declares its C signatures and
raises HydftError on a nonzero status.  It never falls back to anything: if
libhydft.so is missing, importing the binding fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIBPATH = Path(os.environ.get("HYDFT_LIB", Path(__file__).resolve().parent / "libhydft.so"))

HY_OK = 0
HY_ERR_INVALID = 1
HY_ERR_NONPOSITIVE = 2
HY_ERR_OOM = 3
HY_ERR_HIP = 4
HY_ERR_UNKNOWN_KIND = 5
HY_ERR_UNSUPPORTED = 6
HY_ERR_NO_DT = 7
HY_ERR_BUSY = 8

HY_PREC_FP64 = 0
HY_PREC_MIXED = 1

HY_TERM_IDEAL, HY_TERM_TF, HY_TERM_GRADIENT, HY_TERM_HARTREE, HY_TERM_RY = range(5)
HY_CLOSURE_NEWTONIAN, HY_CLOSURE_MAXWELL = 0, 1
HY_VEXT_NONE, HY_VEXT_MODE, HY_VEXT_DRIVEN, HY_VEXT_GAUSSIAN = range(4)


class HydftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hydft error {code}: {msg}")
        self.code = code


class TermDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("n0", C.c_double), ("beta", C.c_double),
        ("e2", C.c_double), ("kappa", C.c_double),
        ("cn", C.c_double), ("alpha0", C.c_double), ("cg", C.c_double),
        ("subtract_mean_field", C.c_int),
        ("ntab", C.c_int),
        ("k_tab", C.POINTER(C.c_double)),
        ("c_tab", C.POINTER(C.c_double)),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("dealias", C.c_int), ("precision", C.c_int),
        ("mass", C.c_double), ("n0", C.c_double),
        ("nterms", C.c_int), ("terms", C.POINTER(TermDesc)),
        ("closure", C.c_int),
        ("eta", C.c_double), ("xi", C.c_double), ("tau", C.c_double),
        ("eta_l", C.c_double),
        ("vext_kind", C.c_int), ("vext_amp", C.c_double), ("vext_mode", C.c_int * 3),
        ("vext_omega", C.c_double), ("vext_width", C.c_double), ("vext_ramp", C.c_double),
        ("dt", C.c_double), ("cfl", C.c_double),
    ]


_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)

_SIGS = {
    "hy_init": [_I],
    "hy_version": [],
    "hy_device_sync": [],
    "hy_grid_create": [_I, _I, _I, _D, _D, _D, _I, _I, C.POINTER(_P)],
    "hy_grid_destroy": [_P],
    "hy_grid_query": [_P, _PI, _PI, _PI, _PD, _PD],
    "hy_grid_mask": [_P, _P],
    "hy_fft_r2c": [_P, _P, _P, _P],
    "hy_fft_c2r": [_P, _P, _P, _P],
    "hy_spec_grad_hat": [_P, _P, _I, _P, _P],
    "hy_spec_grad": [_P, _P, _P, _P],
    "hy_spec_lap": [_P, _P, _P, _P],
    "hy_spec_div": [_P, _P, _P, _P],
    "hy_spec_convolve": [_P, _P, _P, _P, _P],
    "hy_spec_dealias": [_P, _P, _P],
    "hy_spec_mean": [_P, _P, _PD, _P],
    "hy_spec_mean_noalloc": [_P, _P, _PD, _P],
    "hy_spec_mode": [_P, _P, _I, _I, _I, _PD, _P],
    "hy_stress_pi0": [_P, _D, _D, _P, _P, _P],
    "hy_stress_div": [_P, _P, _P, _P],
    "hy_model_create": [C.POINTER(ModelDesc), C.POINTER(_P)],
    "hy_model_destroy": [_P],
    "hy_model_set_state": [_P, _P, _P, _P],
    "hy_model_get_state": [_P, _P, _P, _P],
    "hy_model_set_time": [_P, _D, _I],
    "hy_model_get_time": [_P, _PD, _PD, _PI],
    "hy_model_set_dt": [_P, _D],
    "hy_model_compute_dt": [_P, _PD],
    "hy_model_advance": [_P, _I, _I, _D, _PI, _P, _PI],
    "hy_model_rhs": [_P, _D, _P, _P, _P],
    "hy_model_mu": [_P, _P, _P],
    "hy_model_budgets": [_P, _P],
    "hy_model_probe": [_P, _I, _I, _I, _P],
    "hy_model_kernel": [_P, _D, _PD],
    "hy_fd_eval": [_I, _I, _P, _P],
    "hy_device_alloc": [C.c_ulonglong, C.POINTER(_P)],
    "hy_device_free": [_P],
    "hy_copy_to_device": [_P, _P, C.c_ulonglong],
    "hy_copy_to_host": [_P, _P, C.c_ulonglong],
    "hy_model_profile": [_P, _I],
    "hy_model_profile_classes": [_P, C.c_uint],
    "hy_model_profile_read": [_P, _I, _P, _P, _P],
    "hy_model_graph_build": [_P, _I, _I, _PI, _I],
    "hy_model_graph_launch": [_P, _I],
    "hy_model_graph_probe": [_P, _P],
    "hy_model_idle": [_P, _PI],
    "hy_model_check": [_P],
    "hy_model_device_bytes": [_P, C.POINTER(C.c_ulonglong)],
    "hy_batch_create": [_P, _I, _I, _I, _PI, _I, C.POINTER(_P)],
    "hy_batch_launch": [_P, _I],
    "hy_batch_set_stops": [_P, _P],
    "hy_group_create": [_P, _I, C.POINTER(_P)],
    "hy_group_launch": [_P, _I],
    "hy_group_idle": [_P, _PI],
    "hy_group_sync": [_P],
    "hy_group_destroy": [_P],
    "hy_batch_idle": [_P, _PI],
    "hy_batch_sync": [_P],
    "hy_batch_destroy": [_P],
    "hy_model_probe_ring": [_P, _I, _P],
    "hy_model_kept": [_P, _PI],
    "hy_plan_tuning": [_I],
    "hy_plan_reset": [],
    "hy_plan_list": [_I, _P],
    "hy_fun_create": [_P, _I, _P, C.POINTER(_P)],
    "hy_fun_destroy": [_P],
    "hy_fun_nterms": [_P, _PI],
    "hy_fun_add_mu": [_P, _I, _P, _P, _P, _P],
    "hy_fun_mu": [_P, _P, _P, _P, _P],
    "hy_fun_kernel": [_P, _I, _D, _PD],
    "hy_fun_energy": [_P, _I, _P, _P, _PD, _P],
    "hy_hnc_iterate": [_I, _I, C.c_double, C.c_double, _P, _P, _P, _P, C.c_double, C.c_double, _I, _P, _P, _P, _P],
}


def declared_symbols():
    """Every function name declared in include/hydft.h."""
    import re
    hdr = Path(__file__).resolve().parent.parent / "include" / "hydft.h"
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(hy_\w+)\s*\(", hdr.read_text(), re.M)))


class _Lib:
    def __init__(self, path=LIBPATH):
        if not Path(path).exists():
            raise ImportError(f"libhydft.so not built at {path}; run `python -m hydft_amd.build`")
        self.lib = C.CDLL(str(path))
        self.lib.hy_last_error.restype = C.c_char_p
        self.lib.hy_last_error.argtypes = []
        self.lib.hy_profile_class_name.restype = C.c_char_p
        self.lib.hy_profile_class_name.argtypes = [C.c_int]
        self.missing = []
        for name, args in _SIGS.items():
            try:
                fn = getattr(self.lib, name)
            except AttributeError:
                self.missing.append(name)
                continue
            fn.argtypes = args
            fn.restype = C.c_int

    def call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if name in ("hy_model_profile_read", "hy_plan_list") and rc > 0:
            return rc
        if rc != HY_OK:
            raise HydftError(rc, self.lib.hy_last_error().decode())
        return rc


_lib = None


def lib() -> _Lib:
    global _lib
    if _lib is None:
        _lib = _Lib()
    return _lib


def call(name, *args):
    return lib().call(name, *args)
