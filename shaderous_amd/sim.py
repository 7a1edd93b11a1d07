"""The uniform-grid simulation, mirroring the reference's `Simulation` interface for the hydro path.

The reference's TypeScript API (src/sim.ts:465-627, `Sim` :1938-3882) is kept by name and meaning:
``SimConfig``/``config_for`` (src/sim.ts:48-264, 645-736), ``Sim.create`` (:2283), ``step`` (:3592),
``run_to_end`` (runToEnd :3689-3718), ``read_state``/``set_state`` (:3790, :3031), ``diagnostics``
(:1647-1673), ``flux_mismatch`` (:3724-3787), ``profile`` (:3541-3590), ``destroy`` (:3852). Every
call goes through the C ABI of libshx (include/shx.h); there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib
from ._lib import ShxConfig, check


@dataclass
class SimConfig:
    """The SimConfig fields of the hydro path; defaults are configFor's (src/sim.ts:645-736)."""

    n: tuple = (64, 64, 64)
    ndim: int = 3
    dx: float = 1.0 / 64
    gamma: float = 1.4
    bc_lo: tuple = (0, 0, 0)
    bc_hi: tuple = (0, 0, 0)
    limiter: str = "koren"
    riemann: str = "hllc"
    isentropic: bool = True
    cfl: float = 0.8
    rho_floor: float = 1e-10
    p_floor: float = 1e-14
    eta_dual: float = 1e-3
    shock_thresh: float = 0.5
    t_end: float = math.inf
    fluff: tuple = (0.0, 0.0, 0.0)
    precision: int = 32
    path: str = "fused"
    device: int = 0
    tile: tuple = (0, 0, 0)
    org: tuple = None       # the domain's lower corner; None: -n[0] dx / 2 on every axis (domainLo)

    def to_c(self) -> ShxConfig:
        c = ShxConfig()
        c.n[:] = [int(v) for v in self.n]
        c.ndim = int(self.ndim)
        c.bc_lo[:] = [int(v) for v in self.bc_lo]
        c.bc_hi[:] = [int(v) for v in self.bc_hi]
        c.limiter = _lib.LIMITERS[self.limiter]
        c.riemann = _lib.RIEMANN[self.riemann]
        c.isentropic = int(bool(self.isentropic))
        c.precision = int(self.precision)
        c.path = _lib.PATHS[self.path]
        c.device = int(self.device)
        c.dx = float(self.dx)
        c.gamma = float(self.gamma)
        c.cfl = float(self.cfl)
        c.rho_floor, c.p_floor = float(self.rho_floor), float(self.p_floor)
        c.eta_dual, c.shock_thresh = float(self.eta_dual), float(self.shock_thresh)
        c.t_end = float(self.t_end) if math.isfinite(self.t_end) else 1e30
        c.fluff[:] = [float(v) for v in self.fluff]
        c.tile[:] = [int(v) for v in self.tile]
        org = self.org if self.org is not None else (-0.5 * self.n[0] * self.dx,) * 3
        c.org[:] = [float(v) for v in org]
        return c


def config_for(problem, **over) -> SimConfig:
    """configFor (src/sim.ts:645-736) for a shaderous_amd.problems.Problem, with overrides."""
    cfg = SimConfig(n=tuple(problem.n), ndim=problem.ndim, dx=problem.dx, gamma=problem.gamma,
                    bc_lo=tuple(problem.bc_lo), bc_hi=tuple(problem.bc_hi), limiter=problem.limiter,
                    riemann=problem.riemann, isentropic=problem.isentropic, cfl=problem.cfl,
                    rho_floor=problem.rho_floor, p_floor=problem.p_floor, t_end=problem.t_end)
    return replace(cfg, **over)


@dataclass
class Diagnostics:
    """Sim.diagnostics (src/sim.ts:1647-1673): the 24 channels of hydro/diagnostics.wgsl, every
    cell weighted by its volume, summed in FP64; `raw` holds the channels as the ABI returns them."""

    t: float
    dt: float
    steps: int
    held: bool
    mass: float
    momentum: tuple
    e_kin: float
    e_int: float
    e_grav: float
    e_tot: float
    max_mach: float
    floored: float
    jeans_unresolved: float
    nuclear_heating: float
    com: tuple
    ranges: dict            # (min, max) of rho, T, s, p, speed
    raw: np.ndarray = field(repr=False, default=None)


def diagnostics_from(r, ck) -> Diagnostics:
    """diagnosticsFrom (src/sim.ts:1647-1673): the record from the channels and the clock."""
    names = ("rho", "T", "s", "p", "speed")
    return Diagnostics(t=ck[1] + ck[0], dt=ck[0], steps=int(ck[3] + (1 if ck[0] > 0 else 0)), held=ck[10] != 0,
                       mass=r[0], momentum=tuple(r[1:4]), e_kin=r[4], e_int=r[5], e_grav=r[6],
                       e_tot=r[4] + r[5] + r[6], max_mach=r[7], floored=r[8], jeans_unresolved=r[9],
                       nuclear_heating=r[23], com=tuple(r[20:23] / r[0]) if r[0] else (0.0, 0.0, 0.0),
                       ranges={k: (r[10 + 2 * i], r[11 + 2 * i]) for i, k in enumerate(names)}, raw=r)


class Sim:
    """A uniform-grid hydro run on one GPU (the reference's `Sim`)."""

    def __init__(self, cfg: SimConfig):
        lib = _lib.load()
        self.cfg = cfg
        self.dtype = np.float64 if cfg.precision == 64 else np.float32
        self.ncell = int(np.prod(cfg.n))
        self._h = C.c_void_p()
        c = cfg.to_c()
        check(lib.shx_create(C.byref(c), C.byref(self._h)))

    @classmethod
    def create(cls, cfg: SimConfig) -> "Sim":
        return cls(cfg)

    # -- state --
    def set_state(self, u1, u2, time: float = 0.0, steps: int = 0):
        u1 = np.ascontiguousarray(u1, self.dtype).reshape(-1)
        u2 = np.ascontiguousarray(u2, self.dtype).reshape(-1)
        if u1.size != 4 * self.ncell or u2.size != 4 * self.ncell:
            raise ValueError(f"state must hold {self.ncell} cells of 4 values")
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_set_state_f64(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P), self.ncell, time, steps))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_set_state(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P), self.ncell, time, steps))

    def read_state(self):
        """(u1, u2) as (ncell, 4) arrays (Sim.readState)."""
        u1 = np.empty((self.ncell, 4), self.dtype)
        u2 = np.empty((self.ncell, 4), self.dtype)
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_read_state_f64(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_read_state(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        return u1, u2

    def reset(self, problem):
        """Load a problem's initial conditions (Sim.reset)."""
        u1, u2 = problem.state(self.dtype)
        self.set_state(u1, u2)

    # -- stepping --
    def step(self, count: int = 1) -> int:
        """Enqueue `count` steps (encodeSteps); returns the steps taken."""
        n = C.c_int(0)
        check(_lib.load().shx_step(self._h, int(count), C.byref(n)))
        return n.value

    def sync(self):
        check(_lib.load().shx_sync(self._h))

    def run_to_end(self, batch: int = 16, max_steps: int = 1 << 30) -> int:
        """Sim.runToEnd: step until the clock holds at t_end; the clock's step count then."""
        n = C.c_int(0)
        check(_lib.load().shx_run_to_end(self._h, batch, max_steps, C.byref(n)))
        return n.value

    def clock(self) -> np.ndarray:
        out = np.zeros(11)
        check(_lib.load().shx_read_clock(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def diagnostics(self) -> Diagnostics:
        out = np.zeros(24)
        check(_lib.load().shx_diagnostics(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return diagnostics_from(out, self.clock())

    def write_params(self, cfg: SimConfig = None):
        """Sim.writeParams: the run-time fields of `cfg` (default self.cfg) from the next step on."""
        cfg = cfg or self.cfg
        c = cfg.to_c()
        check(_lib.load().shx_write_params(self._h, C.byref(c)))
        self.cfg = cfg

    def build_hydro(self, cfg: SimConfig = None):
        """Sim.buildHydro: the kernels of `cfg`'s limiter, Riemann solver and reconstruction."""
        cfg = cfg or self.cfg
        c = cfg.to_c()
        check(_lib.load().shx_build_hydro(self._h, C.byref(c)))
        self.cfg = cfg

    def flux_mismatch(self):
        """(mismatched faces, faces compared) (Sim.fluxMismatch)."""
        m, f = C.c_int64(0), C.c_int64(0)
        check(_lib.load().shx_flux_mismatch(self._h, C.byref(m), C.byref(f)))
        return m.value, f.value

    def profile(self, steps: int = 10) -> dict:
        ms = np.zeros(5)
        check(_lib.load().shx_profile(self._h, steps, ms.ctypes.data_as(C.POINTER(C.c_double))))
        return {"rates": ms[0], "clock": ms[1], "predict": ms[2], "update": ms[3], "step": ms[4]}

    def kernel_name(self) -> str:
        buf = C.create_string_buffer(64)
        check(_lib.load().shx_kernel_name(self._h, buf, 64))
        return buf.value.decode()

    def autotune(self, reps: int = 3):
        out = (C.c_int * 3)()
        check(_lib.load().shx_autotune(self._h, reps, out))
        return tuple(out)

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().shx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
