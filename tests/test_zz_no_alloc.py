"""shx_diagnostics_noalloc (the diagnostics over a caller's device state, with the library's
file-scope scratch array instead of an allocation) through the C ABI: on two calls in a row, the
second on a smaller grid, and in both precisions, its 24 channels equal shx_diagnostics' of a sim
holding the same state bit for bit (the same kernels reduce in the same fixed order)."""
import ctypes as C

import numpy as np
import pytest

from shaderous_amd import _lib, problems
from shaderous_amd._lib import ShxConfig, check
from shaderous_amd.sim import Sim, config_for

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def noalloc_diag(h, cfg, u1, u2):
    """Copy (u1, u2) into device buffers this test allocates, run the entry point on the default
    stream, and return its 24 channels."""
    lib = _lib.load()
    fn = lib.shx_diagnostics_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(ShxConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    bufs = []
    try:
        for a in (u1, u2):
            p = C.c_void_p()
            assert h.hipMalloc(C.byref(p), a.nbytes) == 0
            bufs.append(p)
            assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        out = C.c_void_p()
        assert h.hipMalloc(C.byref(out), 24 * 8) == 0
        bufs.append(out)
        c = cfg.to_c()
        check(fn(C.byref(c), bufs[0], bufs[1], out, None))
        assert h.hipDeviceSynchronize() == 0
        res = np.zeros(24)
        assert h.hipMemcpy(res.ctypes.data, out, 24 * 8, D2H) == 0
        return res
    finally:
        for p in bufs:
            h.hipFree(p)


def sim_diag(cfg, u1, u2):
    with Sim(cfg) as sim:
        sim.set_state(u1, u2)
        return sim.diagnostics().raw


@pytest.mark.parametrize("precision", [32, 64])
def test_noalloc_diagnostics_equal_the_sims(precision):
    h = hip()
    dtype = np.float64 if precision == 64 else np.float32
    cases = [problems.turbulence(64, limiter="mc", isentropic=False), problems.kelvin_helmholtz(48)]
    for pb in cases:     # the second call: a smaller grid
        cfg = config_for(pb, precision=precision)
        u1, u2 = pb.state(dtype)
        got = noalloc_diag(h, cfg, np.ascontiguousarray(u1), np.ascontiguousarray(u2))
        want = sim_diag(cfg, u1, u2)
        assert np.array_equal(got, want), (pb.id, got - want)


def test_noalloc_refuses_bad_arguments():
    lib = _lib.load()
    fn = lib.shx_diagnostics_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(ShxConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    c = ShxConfig()
    lib.shx_config_default(C.byref(c), 2048, 2048, 1024)      # 4.3e9 cells: beyond the reduction's range
    dummy = C.c_void_p(256)
    assert fn(C.byref(c), dummy, dummy, dummy, None) == _lib.SHX_EINVAL
    assert b"at most" in lib.shx_last_error()
    assert fn(None, dummy, dummy, dummy, None) == _lib.SHX_EINVAL
