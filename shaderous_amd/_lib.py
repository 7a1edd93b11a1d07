"""ctypes binding of libshx.so (include/shx.h). Loads the in-tree library and fails loudly when
it is missing or cannot see a GPU: there is no CPU fallback on the product path."""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "shx.h")

SHX_OK, SHX_EINVAL, SHX_ENOMEM, SHX_EDEVICE, SHX_ESTATE = 0, -1, -2, -3, -4
LIMITERS = {"koren": 0, "mc": 1, "vanleer": 2, "minmod": 3}
RIEMANN = {"hllc": 0, "hll": 1}
PATHS = {"fused": 0, "plain": 1}
AMR_PATHS = {"fused": 0, "passes": 1, "levels": 2}


class ShxError(RuntimeError):
    """A libshx call returned a failure status (the reference throws Error for the same)."""

    def __init__(self, code, msg):
        super().__init__(f"libshx error {code}: {msg}")
        self.code = code


class ShxConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3), ("ndim", C.c_int), ("bc_lo", C.c_int * 3), ("bc_hi", C.c_int * 3),
        ("limiter", C.c_int), ("riemann", C.c_int), ("isentropic", C.c_int), ("precision", C.c_int),
        ("path", C.c_int), ("device", C.c_int),
        ("dx", C.c_double), ("gamma", C.c_double), ("cfl", C.c_double), ("rho_floor", C.c_double),
        ("p_floor", C.c_double), ("eta_dual", C.c_double), ("shock_thresh", C.c_double), ("t_end", C.c_double),
        ("fluff", C.c_double * 3), ("tile", C.c_int * 3), ("org", C.c_double * 3),
    ]


class ShxAmrLayout(C.Structure):
    _fields_ = [("root", C.c_int * 3), ("path", C.c_int), ("compensated", C.c_int), ("time_bins", C.c_int),
                ("regrid_every", C.c_int), ("max_level", C.c_int), ("refine_above", C.c_double),
                ("coarsen_below", C.c_double)]


_lib = None


def declared_symbols():
    """Every function include/shx.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\**(shx_[a-z0-9_]+)\(", text, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ShxError(SHX_EDEVICE, f"{LIB_PATH} is not built: run `python -m shaderous_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, D, I = C.c_void_p, C.POINTER(C.c_double), C.c_int
    F = C.POINTER(C.c_float)
    sig = {
        "shx_config_default": (None, [C.POINTER(ShxConfig), I, I, I]),
        "shx_create": (I, [C.POINTER(ShxConfig), C.POINTER(P)]),
        "shx_destroy": (I, [P]),
        "shx_set_state": (I, [P, F, F, C.c_size_t, C.c_double, I]),
        "shx_set_state_f64": (I, [P, D, D, C.c_size_t, C.c_double, I]),
        "shx_read_state": (I, [P, F, F]),
        "shx_read_state_f64": (I, [P, D, D]),
        "shx_step": (I, [P, I, C.POINTER(I)]),
        "shx_sync": (I, [P]),
        "shx_run_to_end": (I, [P, I, I, C.POINTER(I)]),
        "shx_read_clock": (I, [P, D]),
        "shx_diagnostics": (I, [P, D]),
        "shx_flux_mismatch": (I, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "shx_profile": (I, [P, I, D]),
        "shx_kernel_name": (I, [P, C.c_char_p, C.c_size_t]),
        "shx_autotune": (I, [P, I, C.POINTER(I)]),
        "shx_write_params": (I, [P, C.POINTER(ShxConfig)]),
        "shx_build_hydro": (I, [P, C.POINTER(ShxConfig)]),
        "shx_readback_begin": (I, [P, I, C.POINTER(P)]),
        "shx_amr_readback_begin": (I, [P, I, C.POINTER(P)]),
        "shx_readback_wait": (I, [P]),
        "shx_readback_data": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(C.c_size_t), C.POINTER(I),
                                  C.POINTER(D), C.POINTER(D)]),
        "shx_readback_free": (I, [P]),
        "shx_amr_create": (I, [C.POINTER(ShxConfig), C.POINTER(ShxAmrLayout), C.POINTER(I), I, C.POINTER(P)]),
        "shx_amr_destroy": (I, [P]),
        "shx_amr_table": (I, [P, C.POINTER(I), C.c_size_t]),
        "shx_amr_set_state": (I, [P, F, F, C.c_size_t, C.c_double, I]),
        "shx_amr_set_state_f64": (I, [P, D, D, C.c_size_t, C.c_double, I]),
        "shx_amr_read_state": (I, [P, F, F]),
        "shx_amr_read_state_f64": (I, [P, D, D]),
        "shx_amr_step": (I, [P, I, C.POINTER(I)]),
        "shx_amr_sync": (I, [P]),
        "shx_amr_read_clock": (I, [P, D]),
        "shx_amr_run_to_end": (I, [P, I, I, C.POINTER(I)]),
        "shx_amr_bins": (I, [P, C.POINTER(I), C.c_size_t, C.POINTER(C.c_int64)]),
        "shx_amr_write_params": (I, [P, C.POINTER(ShxConfig)]),
        "shx_amr_build_hydro": (I, [P, C.POINTER(ShxConfig)]),
        "shx_amr_diagnostics": (I, [P, D]),
        "shx_amr_profile": (I, [P, I, D]),
        "shx_amr_stats": (I, [P, C.POINTER(I)]),
        "shx_amr_flux_mismatch": (I, [P, C.POINTER(C.c_int64), D]),
        "shx_amr_indicator": (I, [P, C.POINTER(C.c_double)]),
        "shx_amr_regrid": (I, [P, C.POINTER(I), I]),
        "shx_amr_adapt": (I, [C.POINTER(I), I, C.POINTER(I), I, C.POINTER(C.c_double), C.c_double, C.c_double, I,
                              C.POINTER(I), C.POINTER(I), I, C.POINTER(I)]),
        "shx_ens_create": (I, [C.POINTER(ShxConfig), I, C.POINTER(I), C.POINTER(P)]),
        "shx_ens_destroy": (I, [P]),
        "shx_ens_set_state": (I, [P, I, F, F, C.c_size_t]),
        "shx_ens_set_state_f64": (I, [P, I, D, D, C.c_size_t]),
        "shx_ens_read_state": (I, [P, I, F, F]),
        "shx_ens_read_state_f64": (I, [P, I, D, D]),
        "shx_ens_step": (I, [P, I]),
        "shx_ens_sync": (I, [P]),
        "shx_ens_read_clock": (I, [P, I, D]),
        "shx_ens_profile": (I, [P, I, D]),
        "shx_ens_autotune": (I, [P, I, C.POINTER(I)]),
        "shx_last_error": (C.c_char_p, []),
        "shx_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != SHX_OK:
        raise ShxError(code, load().shx_last_error().decode())
    return code
