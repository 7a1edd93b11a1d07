"""Ensembles of independent 2D grids advanced together (include/shx.h, shx_ens_*).

The reference runs one Sim per device (src/sim.ts); an Ensemble holds `count` 2D uniform grids of
their own sizes, each with its own clock and dt, and steps all of them with one persistent
launch per step. Each instance evolves as a Sim of the same configuration would (the plain
path's arithmetic). There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .sim import SimConfig


class Ensemble:
    """`cfg` gives what the instances share (limiter, solver, gamma, cfl, boundaries, floors,
    precision); `sizes` is a list of (nx, ny), each instance on the unit box."""

    def __init__(self, cfg: SimConfig, sizes):
        lib = _lib.load()
        self.cfg = cfg
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self.count = len(self.sizes)
        self.dtype = np.float64 if cfg.precision == 64 else np.float32
        arr = (C.c_int * (2 * self.count))(*[v for ab in self.sizes for v in ab])
        self._h = C.c_void_p()
        c = cfg.to_c()
        check(lib.shx_ens_create(C.byref(c), self.count, arr, C.byref(self._h)))

    def set_state(self, i, u1, u2):
        n = self.sizes[i][0] * self.sizes[i][1]
        u1 = np.ascontiguousarray(u1, self.dtype).reshape(-1)
        u2 = np.ascontiguousarray(u2, self.dtype).reshape(-1)
        if u1.size != 4 * n or u2.size != 4 * n:
            raise ValueError(f"instance {i} holds {n} cells of 4 values")
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_ens_set_state_f64(self._h, i, u1.ctypes.data_as(P), u2.ctypes.data_as(P), n))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_ens_set_state(self._h, i, u1.ctypes.data_as(P), u2.ctypes.data_as(P), n))

    def read_state(self, i):
        n = self.sizes[i][0] * self.sizes[i][1]
        u1 = np.empty((n, 4), self.dtype)
        u2 = np.empty((n, 4), self.dtype)
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_ens_read_state_f64(self._h, i, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_ens_read_state(self._h, i, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        return u1, u2

    def step(self, count: int = 1):
        check(_lib.load().shx_ens_step(self._h, int(count)))

    def sync(self):
        check(_lib.load().shx_ens_sync(self._h))

    def clock(self, i) -> np.ndarray:
        out = np.zeros(11)
        check(_lib.load().shx_ens_read_clock(self._h, i, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def profile(self, steps: int = 5) -> dict:
        ms = np.zeros(3)
        check(_lib.load().shx_ens_profile(self._h, steps, ms.ctypes.data_as(C.POINTER(C.c_double))))
        return {"clock": ms[0], "fused": ms[1], "step": ms[2]}

    def autotune(self, reps: int = 3):
        out = (C.c_int * 2)()
        check(_lib.load().shx_ens_autotune(self._h, reps, out))
        return tuple(out)

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().shx_ens_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
