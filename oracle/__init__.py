"""The oracle: CPU restatements of the reference's hydro path, as TEST INFRASTRUCTURE.

Only the test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
this package. It wraps ``oracle/_build/liboracle.so`` (built by ``oracle/Makefile``):

- ``wgsl32`` / ``wgsl64``: the WGSL GPU scheme (src/hydro/*.wgsl) in FP32 and FP64;
- ``RefHydro``: the reference's float64 CPU scheme (src/reference/hydro.ts:172-458);
- ``exact_riemann``: Toro's exact solver (src/reference/riemann_exact.ts:9-71);
- ``AmrTable`` / ``Wgsl.amr_step`` / ``Wgsl.amr_prolong``: the block table of a fixed block tree
  (src/amr/tree.ts) and one block-AMR step (the src/amr WGSL passes), 3D.

Parity status: every restated function is pinned by tests/test_oracle.py (known answers,
closed forms, invariants, and cross-checks between the two independent restatements).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

LIM = {"koren": 0, "mc": 1, "vanleer": 2, "minmod": 3}
RS = {"hllc": 0, "hll": 1}


class OrcParams(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3), ("bc_lo", C.c_int * 3), ("bc_hi", C.c_int * 3),
        ("ndim", C.c_int), ("limiter", C.c_int), ("riemann", C.c_int), ("isentropic", C.c_int),
        ("gamma", C.c_double), ("dx", C.c_double), ("cfl", C.c_double),
        ("rho_floor", C.c_double), ("p_floor", C.c_double), ("eta_dual", C.c_double),
        ("shock_thresh", C.c_double), ("t_end", C.c_double), ("dt_max", C.c_double),
        ("fluff", C.c_double * 3),
    ]


class OrcAmr(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("root", C.c_int * 3), ("bc_lo", C.c_int * 3), ("bc_hi", C.c_int * 3),
                ("blk", C.POINTER(C.c_int)), ("tab", C.POINTER(C.c_int)), ("ndim", C.c_int)]


BT_STRIDE = 32
BS = 16
PAD = 20


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        for pfx, ct in (("orc_f32_", C.c_float), ("orc_f64_", C.c_double)):
            P = C.POINTER(ct)
            getattr(_lib, pfx + "step").argtypes = [C.POINTER(OrcParams), P, P, P, P]
            getattr(_lib, pfx + "max_rate").argtypes = [C.POINTER(OrcParams), P, P]
            getattr(_lib, pfx + "max_rate").restype = ct
            getattr(_lib, pfx + "clock").argtypes = [C.POINTER(OrcParams), P, ct]
            getattr(_lib, pfx + "predict").argtypes = [C.POINTER(OrcParams), P, P, ct, P, P]
            getattr(_lib, pfx + "update").argtypes = [C.POINTER(OrcParams), P, P, P, P, ct, P, P]
            getattr(_lib, pfx + "face_mass").argtypes = [C.POINTER(OrcParams), P, P, P, P, P]
            getattr(_lib, pfx + "edge4").argtypes = [C.c_int, P, P, P, P]
            getattr(_lib, pfx + "prim").argtypes = [C.POINTER(OrcParams), P, P, P, P]
            getattr(_lib, pfx + "riemann").argtypes = [C.POINTER(OrcParams), P, P, ct, P]
            getattr(_lib, pfx + "predictor").argtypes = [C.POINTER(OrcParams), P, ct, P]
            getattr(_lib, pfx + "amr_step").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), C.c_double, C.c_int,
                                                        P, P, P]
            getattr(_lib, pfx + "amr_assemble").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), C.c_double,
                                                            P, P, P]
            getattr(_lib, pfx + "amr_assemble").restype = ct
            getattr(_lib, pfx + "amr_flag").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), C.c_double, P, P, P]
            getattr(_lib, pfx + "amr_prolong").argtypes = [C.POINTER(OrcParams), P, C.POINTER(C.c_int),
                                                           C.POINTER(C.c_int), P]
            getattr(_lib, pfx + "amr_regrid").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), P, P, C.c_int,
                                                          C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                          P, P]
            getattr(_lib, pfx + "amr_bins").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), C.c_int, P, ct,
                                                        C.POINTER(C.c_int)]
            getattr(_lib, pfx + "amr_cycle").argtypes = [C.POINTER(OrcParams), C.POINTER(OrcAmr), C.c_double, C.c_int,
                                                         C.c_int, C.POINTER(C.c_int), P, P, P, C.POINTER(C.c_int),
                                                         C.POINTER(C.c_longlong)]
        I3 = C.c_int * 3
        _lib.orc_amr_build.argtypes = [C.POINTER(OrcAmr), C.POINTER(C.c_int), C.c_int, I3, I3, I3]
        _lib.orc_amr_build_nd.argtypes = [C.POINTER(OrcAmr), C.POINTER(C.c_int), C.c_int, I3, I3, I3, C.c_int]
        _lib.orc_amr_free.argtypes = [C.POINTER(OrcAmr)]
        _lib.orc_ref_create.argtypes = [C.POINTER(OrcParams)]
        _lib.orc_ref_create.restype = C.c_void_p
        _lib.orc_ref_destroy.argtypes = [C.c_void_p]
        _lib.orc_ref_U.argtypes = [C.c_void_p, C.c_int]
        _lib.orc_ref_U.restype = C.POINTER(C.c_double)
        _lib.orc_ref_t.argtypes = [C.c_void_p]
        _lib.orc_ref_t.restype = C.c_double
        _lib.orc_ref_steps.argtypes = [C.c_void_p]
        _lib.orc_ref_step.argtypes = [C.c_void_p, C.c_double]
        _lib.orc_ref_step.restype = C.c_double
        _lib.orc_ref_step_slab.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.orc_ref_step_slab.restype = C.c_double
        _lib.orc_ref_set_threads.argtypes = [C.c_int]
        _lib.orc_exact_riemann.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                           C.c_double, C.POINTER(C.c_double)]
    return _lib


def params(n, *, ndim=None, dx=None, gamma=1.4, cfl=0.8, bc_lo=(0, 0, 0), bc_hi=(0, 0, 0),
           limiter="koren", riemann="hllc", isentropic=False, rho_floor=1e-10, p_floor=1e-14,
           eta_dual=1e-3, shock_thresh=0.5, t_end=1e30, dt_max=1e30, fluff=(0.0, 0.0, 0.0)) -> OrcParams:
    n = tuple(int(v) for v in n)
    p = OrcParams()
    p.n[:] = n
    p.bc_lo[:] = [int(v) for v in bc_lo]
    p.bc_hi[:] = [int(v) for v in bc_hi]
    p.ndim = ndim if ndim is not None else (2 if n[2] == 1 else 3)
    p.limiter = LIM[limiter] if isinstance(limiter, str) else int(limiter)
    p.riemann = RS[riemann] if isinstance(riemann, str) else int(riemann)
    p.isentropic = int(bool(isentropic))
    p.gamma = gamma
    p.dx = dx if dx is not None else 1.0 / n[0]
    p.cfl = cfl
    p.rho_floor, p.p_floor, p.eta_dual, p.shock_thresh = rho_floor, p_floor, eta_dual, shock_thresh
    p.t_end, p.dt_max = t_end, dt_max
    p.fluff[:] = list(fluff)
    return p


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Wgsl:
    """The WGSL restatement at one precision (np.float32 or np.float64)."""

    def __init__(self, dtype):
        self.dtype = np.dtype(dtype)
        self.ct = C.c_float if self.dtype == np.float32 else C.c_double
        self.pfx = "orc_f32_" if self.dtype == np.float32 else "orc_f64_"

    def fn(self, name):
        return getattr(lib(), self.pfx + name)

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def step(self, prm, u1, u2, clk):
        """One step in place. u1, u2: (ncell, 4) arrays; clk: 4-array (dt, t, tc, steps)."""
        nc = u1.shape[0]
        work = np.empty(24 * nc, dtype=self.dtype)
        self.fn("step")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct), _ptr(clk, self.ct), _ptr(work, self.ct))

    def max_rate(self, prm, u1, u2):
        return self.fn("max_rate")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct))

    def clock(self, prm, clk, rate):
        self.fn("clock")(C.byref(prm), _ptr(clk, self.ct), rate)

    def predict(self, prm, u1, u2, dt):
        nc = u1.shape[0]
        pa = np.empty((nc, 4), self.dtype)
        pb = np.empty((nc, 4), self.dtype)
        self.fn("predict")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct), dt, _ptr(pa, self.ct), _ptr(pb, self.ct))
        return pa, pb

    def update(self, prm, u1, u2, pa, pb, dt):
        o1 = np.empty_like(u1)
        o2 = np.empty_like(u2)
        self.fn("update")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct), _ptr(pa, self.ct), _ptr(pb, self.ct), dt,
                          _ptr(o1, self.ct), _ptr(o2, self.ct))
        return o1, o2

    def face_mass(self, prm, u1, u2, pa, pb):
        out = np.empty((u1.shape[0], 6), self.dtype)
        self.fn("face_mass")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct), _ptr(pa, self.ct), _ptr(pb, self.ct),
                             _ptr(out, self.ct))
        return out

    def edge4(self, limiter, far, near, scale):
        far, near, scale = (self.arr(np.broadcast_to(x, (4,))) for x in (far, near, scale))
        out = np.empty(4, self.dtype)
        lim = LIM[limiter] if isinstance(limiter, str) else limiter
        self.fn("edge4")(lim, _ptr(far, self.ct), _ptr(near, self.ct), _ptr(scale, self.ct), _ptr(out, self.ct))
        return out

    def riemann(self, prm, l, r, w_hll=0.0):
        l, r = self.arr(l), self.arr(r)
        out = np.empty(8, self.dtype)
        self.fn("riemann")(C.byref(prm), _ptr(l, self.ct), _ptr(r, self.ct), w_hll, _ptr(out, self.ct))
        return out

    def prim(self, prm, u1, u2):
        u1, u2 = self.arr(u1), self.arr(u2)
        a = np.empty(4, self.dtype)
        b = np.empty(4, self.dtype)
        self.fn("prim")(C.byref(prm), _ptr(u1, self.ct), _ptr(u2, self.ct), _ptr(a, self.ct), _ptr(b, self.ct))
        return a, b

    def predictor(self, prm, stencil, dt):
        st = self.arr(stencil).reshape(-1)
        out = np.empty(8, self.dtype)
        self.fn("predictor")(C.byref(prm), _ptr(st, self.ct), dt, _ptr(out, self.ct))
        return out

    def amr_step(self, prm, table, dx0, u1, u2, clk, compensated=True):
        """One block-AMR step in place. u1, u2: (nblocks * 16^ndim, 4) leaf interiors, block by
        block, x fastest; clk: 4-array (dt, t, tc, steps)."""
        assert u1.dtype == self.dtype and u1.flags.c_contiguous and u1.shape == (table.nblocks * BS ** table.ndim, 4)
        assert u2.dtype == self.dtype and u2.flags.c_contiguous and u2.shape == u1.shape
        self.fn("amr_step")(C.byref(prm), C.byref(table.t), dx0, int(bool(compensated)), _ptr(u1, self.ct),
                            _ptr(u2, self.ct), _ptr(clk, self.ct))

    def amr_bins(self, prm, table, kmax, brate, dt):
        """The bins of a cycle (per block) from each block's largest rate / dx and the cycle's dt."""
        br = self.arr(brate)
        assert br.shape == (table.nblocks,)
        out = np.zeros(table.nblocks, np.int32)
        self.fn("amr_bins")(C.byref(prm), C.byref(table.t), int(kmax), _ptr(br, self.ct), dt, _ptr(out, C.c_int))
        return out

    def amr_cycle(self, prm, table, dx0, u1, u2, clk, kmax, bins=None, compensated=True):
        """One cycle of 2^kmax substeps in place (per-block timesteps). bins: fixed bins per block
        (None: the cycle's own). Returns (the bins used, block updates made, leaves << kmax)."""
        assert u1.dtype == self.dtype and u1.flags.c_contiguous and u1.shape == (table.nblocks * BS ** table.ndim, 4)
        assert u2.dtype == self.dtype and u2.flags.c_contiguous and u2.shape == u1.shape
        given = None
        if bins is not None:
            given = np.ascontiguousarray(bins, np.int32)
            assert given.shape == (table.nblocks,)
        out = np.zeros(table.nblocks, np.int32)
        work = (C.c_longlong * 2)()
        self.fn("amr_cycle")(C.byref(prm), C.byref(table.t), dx0, int(bool(compensated)), int(kmax),
                             _ptr(given, C.c_int) if given is not None else None, _ptr(u1, self.ct), _ptr(u2, self.ct),
                             _ptr(clk, self.ct), _ptr(out, C.c_int), work)
        return out, int(work[0]), int(work[1])

    def amr_assemble(self, prm, table, dx0, u1, u2):
        """The assemble pass: (W of shape (nblocks, 20, 20, 20, 8) [z, y, x], in 2D (nblocks, 1, 20,
        20, 8); largest leaf rate/dx)."""
        W = np.zeros((table.nblocks, PAD if table.ndim == 3 else 1, PAD, PAD, 8), self.dtype)
        r = self.fn("amr_assemble")(C.byref(prm), C.byref(table.t), dx0, _ptr(self.arr(u1), self.ct),
                                    _ptr(self.arr(u2), self.ct), _ptr(W, self.ct))
        return W, r

    def amr_flag(self, prm, table, dx0, u1, u2):
        """The refinement flag (src/amr/flag.wgsl:66-155): per block the largest Loehner estimate
        of density and pressure over the leaf's cells, 0 for parents."""
        out = np.zeros(table.nblocks, self.dtype)
        self.fn("amr_flag")(C.byref(prm), C.byref(table.t), dx0, _ptr(self.arr(u1), self.ct),
                            _ptr(self.arr(u2), self.ct), _ptr(out, self.ct))
        return out

    def amr_prolong(self, prm, stencil, ok, par):
        """The fine child of parity par (3 ints of 0/1) of a coarse cell, from its 7-cell stencil
        (centre, -x, +x, -y, +y, -z, +z) of conserved states (7, 8); ok drops missing cells."""
        st = self.arr(stencil).reshape(-1)
        okc = (C.c_int * 7)(*[int(bool(v)) for v in ok])
        pc = (C.c_int * 3)(*[int(v) for v in par])
        out = np.empty(8, self.dtype)
        self.fn("amr_prolong")(C.byref(prm), _ptr(st, self.ct), okc, pc, _ptr(out, self.ct))
        return out

    def amr_regrid(self, prm, table, u1, u2, kind, src, oct):
        """Re-mesh: the new tree's leaf interiors from the old tree's (table, u1, u2), per new slot
        kind (0 none, 1 copy, 2 prolong, 3 restrict), source old slot and child octant. Returns
        (o1, o2) of len(kind) * 16^ndim rows (zero where kind is 0)."""
        kind, src, oct = (np.ascontiguousarray(a, np.int32) for a in (kind, src, oct))
        n = len(kind)
        o1 = np.zeros((n * BS ** table.ndim, 4), self.dtype)
        o2 = np.zeros((n * BS ** table.ndim, 4), self.dtype)
        self.fn("amr_regrid")(C.byref(prm), C.byref(table.t), _ptr(self.arr(u1), self.ct), _ptr(self.arr(u2), self.ct),
                              n, _ptr(kind, C.c_int), _ptr(src, C.c_int), _ptr(oct, C.c_int), _ptr(o1, self.ct),
                              _ptr(o2, self.ct))
        return o1, o2

    def run(self, prm, u1, u2, steps, clk=None):
        """Advance `steps` steps in place; returns the clock."""
        if clk is None:
            clk = np.zeros(4, self.dtype)
        for _ in range(steps):
            self.step(prm, u1, u2, clk)
        return clk


wgsl32 = Wgsl(np.float32)
wgsl64 = Wgsl(np.float64)


class RefHydro:
    """The float64 RefHydro restatement; U is (7, ncell): rho, mx, my, mz, E, e_int, rho*s."""

    def __init__(self, prm: OrcParams):
        self.prm = prm
        self.h = lib().orc_ref_create(C.byref(prm))
        self.ncell = prm.n[0] * prm.n[1] * prm.n[2]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ref_destroy(self.h)
            self.h = None

    def U(self, comp):
        p = lib().orc_ref_U(self.h, comp)
        return np.ctypeslib.as_array(p, shape=(self.ncell,))

    def set_cons(self, u1, u2):
        """From the GPU layout (ncell, 4) x 2."""
        for c in range(4):
            self.U(c)[:] = u1[:, c]
        self.U(4)[:] = u2[:, 0]
        self.U(5)[:] = u2[:, 1]
        self.U(6)[:] = u2[:, 2]

    def cons(self):
        u1 = np.stack([self.U(c) for c in range(4)], axis=1)
        u2 = np.stack([self.U(4), self.U(5), self.U(6), np.zeros(self.ncell)], axis=1)
        return u1, u2

    @property
    def t(self):
        return lib().orc_ref_t(self.h)

    @property
    def steps(self):
        return lib().orc_ref_steps(self.h)

    def step(self, t_end=float("inf")):
        return lib().orc_ref_step(self.h, t_end)

    def step_slab(self, k0, k1):
        return lib().orc_ref_step_slab(self.h, k0, k1)

    def run(self, t_end, max_steps=100000):
        """RefHydro.run (hydro.ts:455-457)."""
        while self.t < t_end * (1 - 1e-12) and self.steps < max_steps:
            self.step(t_end)

    @staticmethod
    def set_threads(n):
        lib().orc_ref_set_threads(int(n))


class AmrTable:
    """The block table of a fixed tree (src/amr/tree.ts:642-675): blocks (nblocks, 4) of
    (level, ix, iy, iz), parents and leaves; boundary codes per axis (0 periodic); ndim 2 for a
    quadtree of 16^2-cell blocks (root[2] = 1, every iz 0)."""

    def __init__(self, blocks, root, bc_lo=(0, 0, 0), bc_hi=(0, 0, 0), ndim=3):
        self.blocks = np.ascontiguousarray(blocks, dtype=np.int32)
        self.nblocks = len(self.blocks)
        self.ndim = ndim
        self.t = OrcAmr()
        rc = lib().orc_amr_build_nd(C.byref(self.t), _ptr(self.blocks, C.c_int), self.nblocks, (C.c_int * 3)(*root),
                                    (C.c_int * 3)(*bc_lo), (C.c_int * 3)(*bc_hi), ndim)
        if rc != 0:
            lib().orc_amr_free(C.byref(self.t))
            raise ValueError(f"block {-1 - rc} breaks 2:1 balance")

    def table(self):
        return np.ctypeslib.as_array(self.t.tab, shape=(self.nblocks, BT_STRIDE)).copy()

    def __del__(self):
        if getattr(self, "t", None) is not None and self.t.tab:
            lib().orc_amr_free(C.byref(self.t))


def exact_riemann(L, R, gamma, xi):
    """Sample Toro's exact solution at xi = x/t: returns (rho, u, p)."""
    Lc = (C.c_double * 3)(*L)
    Rc = (C.c_double * 3)(*R)
    out = (C.c_double * 3)()
    lib().orc_exact_riemann(Lc, Rc, gamma, xi, out)
    return tuple(out)
