"""The C-ABI library loads and exports every symbol include/shx.h declares (CPU only: no compute)."""
import ctypes
import os
import subprocess

import pytest

from shaderous_amd import _lib


def test_header_declares_the_entry_points():
    syms = _lib.declared_symbols()
    for s in ("shx_create", "shx_step", "shx_read_state", "shx_set_state", "shx_run_to_end", "shx_destroy",
              "shx_last_error", "shx_diagnostics", "shx_flux_mismatch", "shx_profile"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libshx.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in _lib.declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports_errors_without_a_gpu_call():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libshx.so not built")
    lib = _lib.load()
    assert isinstance(lib.shx_last_error(), bytes)
    cfg = _lib.ShxConfig()
    lib.shx_config_default(ctypes.byref(cfg), 32, 32, 1)
    assert cfg.ndim == 2 and cfg.limiter == 0 and abs(cfg.cfl - 0.8) < 1e-12
    # an invalid configuration is refused before any device call
    cfg.n[0] = 0
    h = ctypes.c_void_p()
    assert lib.shx_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.SHX_EINVAL
    assert b"n[0]" in lib.shx_last_error()


def test_ensemble_refuses_boundaries_its_kernel_cannot_match():
    """shx_ens_create refuses outflow-only faces and an axis periodic on one face only (before any
    device call): the ensemble's tile kernel matches the plain path only on periodic, outflow and
    reflecting boundaries."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libshx.so not built")
    lib = _lib.load()
    sizes = (ctypes.c_int * 2)(16, 16)
    for lo, hi, what in [((3, 1, 0), (1, 1, 0), b"outflow-only"), ((1, 0, 0), (1, 1, 0), b"one face only")]:
        cfg = _lib.ShxConfig()
        lib.shx_config_default(ctypes.byref(cfg), 16, 16, 1)
        cfg.bc_lo[:] = list(lo)
        cfg.bc_hi[:] = list(hi)
        h = ctypes.c_void_p()
        assert lib.shx_ens_create(ctypes.byref(cfg), 1, sizes, ctypes.byref(h)) == _lib.SHX_EINVAL
        assert what in lib.shx_last_error()
