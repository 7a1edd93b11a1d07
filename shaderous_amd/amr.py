"""The block-AMR simulation, mirroring the reference's `AmrSim` for the hydro path.

The reference's AmrSim (src/amr/amr_sim.ts) runs the same hydro step on 16^3-cell blocks of a
BlockTree (src/amr/tree.ts), or on 16^2-cell blocks of a quadtree in 2D (cfg.ndim 2): ``create`` (:1924-1992, with uploadTree :2566-2625), ``step`` /
encodeSteps (:2709-2947), ``readLeaves``, ``diagnostics``, re-meshing (the flagging pass
:3029-3060 and applyRegrid :2648-2697), ``destroy``. The tree is given as the list of its blocks,
parents and leaves, as (level, ix, iy, iz); ``regrid`` replaces it by one within a level of it
everywhere and ``adapt`` chooses that tree from the refinement indicator. Every call goes through
the C ABI of libshx (include/shx.h, shx_amr_*); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace

import numpy as np

from . import _lib
from ._lib import ShxAmrLayout, check
from .sim import Diagnostics, SimConfig, diagnostics_from

BS = 16
CELLS = BS ** 3


def _level0(root, ndim):
    """The level-0 grid's cells per axis: 16 per root block (one layer along z in 2D)."""
    return (BS * root[0], BS * root[1], 1 if ndim == 2 else BS * root[2])


def adapt_tree(blocks, root, indicator, refine_above, coarsen_below, max_level, periodic=(True, True, True),
               ndim=3):
    """BlockTree.adapt (src/amr/tree.ts:489-641) in libshx (shx_amr_adapt, host code, no GPU):
    the next tree from `blocks` and one indicator per block, as amr_tree.adapt_tree."""
    lib = _lib.load()
    blocks = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1, 4)
    ind = np.ascontiguousarray(np.nan_to_num(np.asarray(indicator, np.float64), nan=0.0))
    r = (C.c_int * 3)(*[int(v) for v in root])
    per = (C.c_int * 3)(*[int(bool(v)) for v in periodic])
    n = C.c_int(0)
    cap = max(4 * len(blocks), 64)
    while True:
        out = np.zeros((cap, 4), np.int32)
        rc = lib.shx_amr_adapt(blocks.ctypes.data_as(C.POINTER(C.c_int)), len(blocks), r, int(ndim),
                               ind.ctypes.data_as(C.POINTER(C.c_double)), float(refine_above), float(coarsen_below),
                               int(max_level), per, out.ctypes.data_as(C.POINTER(C.c_int)), cap, C.byref(n))
        if rc == 0:
            return out[:n.value].copy()
        if n.value <= cap:
            check(rc)
        cap = n.value


class AmrSim:
    """A block-AMR hydro run on one GPU. `cfg` describes the level-0 grid (n = 16 x root blocks,
    dx = level-0 cell size; cfg.ndim 2 for a quadtree, root[2] = 1); `blocks` is (nblocks, 4) of
    (level, ix, iy, iz). regrid_every = N > 0 re-meshes every N steps inside step() (the
    indicator, adapt_tree with refine_above / coarsen_below / max_level, regrid), as the
    reference's AmrSim does every REGRID_EVERY = 8 steps; self.blocks follows the tree."""

    def __init__(self, cfg: SimConfig, blocks, root, path: str = "fused", compensated: bool = True,
                 time_bins: int = 0, regrid_every: int = 0, max_level: int = 0, refine_above: float = 0.8,
                 coarsen_below: float = 0.2):
        lib = _lib.load()
        root = tuple(int(r) for r in root)
        self.ndim = 2 if cfg.ndim == 2 else 3
        self.cells = BS ** self.ndim
        self.cfg = replace(cfg, n=_level0(root, self.ndim), ndim=self.ndim)
        self.root = root
        self.blocks = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1, 4)
        self.nblocks = len(self.blocks)
        self.ncell = self.nblocks * self.cells
        self.dtype = np.float64 if cfg.precision == 64 else np.float32
        lay = ShxAmrLayout()
        lay.root[:] = list(root)
        lay.path = _lib.AMR_PATHS[path]
        lay.compensated = int(bool(compensated))
        lay.time_bins = int(time_bins)
        lay.regrid_every = int(regrid_every)
        lay.max_level = int(max_level)
        lay.refine_above = float(refine_above)
        lay.coarsen_below = float(coarsen_below)
        self.time_bins = int(time_bins)
        self.regrid_every = int(regrid_every)
        self._h = C.c_void_p()
        c = self.cfg.to_c()
        check(lib.shx_amr_create(C.byref(c), C.byref(lay), self.blocks.ctypes.data_as(C.POINTER(C.c_int)),
                                 self.nblocks, C.byref(self._h)))

    # -- tree --
    def table(self) -> np.ndarray:
        out = np.empty((self.nblocks, 32), np.int32)
        check(_lib.load().shx_amr_table(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), out.size))
        return out

    def leaf_mask(self) -> np.ndarray:
        return (self.table()[:, 4] & 2) != 0

    # -- re-meshing --
    def indicator(self) -> np.ndarray:
        """The refinement indicator (amr/flag.wgsl:66-155, gas cells): per leaf the largest
        Loehner estimate of density and pressure over its cells, ghosts included (near 1 at a
        discontinuity); NaN for parents."""
        out = np.empty(self.nblocks, np.float64)
        check(_lib.load().shx_amr_indicator(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def regrid(self, blocks):
        """Replace the tree by `blocks` (within one level of the current tree everywhere): leaves
        copied, prolonged from their parent or restricted from their children; the clock carries
        on."""
        blocks = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1, 4)
        check(_lib.load().shx_amr_regrid(self._h, blocks.ctypes.data_as(C.POINTER(C.c_int)), len(blocks)))
        self.blocks = blocks
        self.nblocks = len(blocks)
        self.ncell = self.nblocks * self.cells

    def adapt(self, refine_above: float, coarsen_below: float, max_level: int, periodic=None) -> bool:
        """One re-mesh from the indicator (amr_tree.adapt_tree): split leaves above
        `refine_above`, merge parents of leaves all below `coarsen_below`. Returns whether the
        tree changed. `periodic` (per axis) defaults to the run's boundaries: an axis wraps for
        the 2:1 balance only where both its faces are periodic."""
        from . import amr_tree
        if periodic is None:
            periodic = tuple(self.cfg.bc_lo[a] == 0 and self.cfg.bc_hi[a] == 0 for a in range(3))
        new = amr_tree.adapt_tree(self.blocks, self.root, self.indicator(), refine_above, coarsen_below, max_level,
                                  periodic, ndim=self.ndim)
        if len(new) == len(self.blocks) and {tuple(b) for b in new.tolist()} == {tuple(b) for b in self.blocks.tolist()}:
            return False
        self.regrid(new)
        return True

    def stats(self) -> dict:
        out = (C.c_int * 5)()
        check(_lib.load().shx_amr_stats(self._h, out))
        return {"blocks": out[0], "leaves": out[1], "match_leaves": out[2], "levels": out[3],
                "launches_per_step": out[4]}

    # -- state --
    def set_state(self, u1, u2, time: float = 0.0, steps: int = 0):
        u1 = np.ascontiguousarray(u1, self.dtype).reshape(-1)
        u2 = np.ascontiguousarray(u2, self.dtype).reshape(-1)
        if u1.size != 4 * self.ncell or u2.size != 4 * self.ncell:
            raise ValueError(f"state must hold {self.nblocks} blocks x {self.cells} cells of 4 values")
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_amr_set_state_f64(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P), self.ncell, time,
                                            steps))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_amr_set_state(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P), self.ncell, time, steps))

    def read_state(self):
        """(u1, u2) as (nblocks * 16^ndim, 4) arrays, block by block; parents' cells are unused."""
        u1 = np.empty((self.ncell, 4), self.dtype)
        u2 = np.empty((self.ncell, 4), self.dtype)
        lib = _lib.load()
        if self.dtype == np.float64:
            P = C.POINTER(C.c_double)
            check(lib.shx_amr_read_state_f64(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        else:
            P = C.POINTER(C.c_float)
            check(lib.shx_amr_read_state(self._h, u1.ctypes.data_as(P), u2.ctypes.data_as(P)))
        return u1, u2

    def read_leaves(self, box_lo: float = -0.5):
        """AmrSim.readLeaves: the leaves' cells as a dict of x, y, z (centres), dx, u1, u2."""
        u1, u2 = self.read_state()
        leaf = self.leaf_mask()
        dx0 = self.cfg.dx
        xs, ys, zs, dxs, sel = [], [], [], [], []
        ax = np.arange(BS) + 0.5
        nc = self.cells
        for i in np.flatnonzero(leaf):
            L, bx, by, bz = (int(v) for v in self.blocks[i])
            h = dx0 / (1 << L)
            zax = (ax + BS * bz) * h + box_lo if self.ndim == 3 else np.zeros(1)
            z, y, x = np.meshgrid(zax, (ax + BS * by) * h + box_lo, (ax + BS * bx) * h + box_lo, indexing="ij")
            xs.append(x.ravel()); ys.append(y.ravel()); zs.append(z.ravel())
            dxs.append(np.full(nc, h))
            sel.append(np.arange(i * nc, (i + 1) * nc))
        sel = np.concatenate(sel)
        return {"x": np.concatenate(xs), "y": np.concatenate(ys), "z": np.concatenate(zs), "dx": np.concatenate(dxs),
                "u1": u1[sel], "u2": u2[sel], "count": sel.size}

    # -- stepping --
    def step(self, count: int = 1) -> int:
        """Enqueue `count` steps (encodeSteps); returns the steps taken. With regrid_every the
        tree may change on the way: self.blocks is read back afterwards."""
        n = C.c_int(0)
        check(_lib.load().shx_amr_step(self._h, int(count), C.byref(n)))
        if self.regrid_every > 0:
            self._refresh_tree()
        return n.value

    def _refresh_tree(self):
        nb = self.stats()["blocks"]
        if nb != self.nblocks:
            self.nblocks = nb
            self.ncell = nb * self.cells
        self.blocks = np.ascontiguousarray(self.table()[:, :4])

    def sync(self):
        check(_lib.load().shx_amr_sync(self._h))

    def clock(self) -> np.ndarray:
        out = np.zeros(11)
        check(_lib.load().shx_amr_read_clock(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def run_to(self, t_end: float, max_steps: int = 100000, batch: int = 16) -> int:
        """Step until the clock holds (dt = 0 at t_end, set through cfg.t_end) or max_steps."""
        taken = 0
        while taken < max_steps:
            n = min(batch, max_steps - taken)
            self.step(n)
            taken += n
            ck = self.clock()
            if not np.isfinite(ck[0]) or not np.isfinite(ck[1]):
                raise _lib.ShxError(_lib.SHX_ESTATE, f"the clock went non-finite (t = {ck[1]}, dt = {ck[0]})")
            if ck[0] == 0:
                break
        return taken

    def bins(self):
        """Per-block timesteps: (the bin of every slot in the last cycle, {cycles, updates, single}):
        block updates made (deposits included) and what single steps would have made."""
        out = np.zeros(self.nblocks, np.int32)
        w = (C.c_int64 * 3)()
        check(_lib.load().shx_amr_bins(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), self.nblocks, w))
        return out, {"cycles": int(w[0]), "updates": int(w[1]), "single": int(w[2])}

    def run_to_end(self, batch: int = 16, max_steps: int = 1 << 30) -> int:
        """AmrSim.runToEnd: step until the clock holds at cfg.t_end; the clock's step count then."""
        n = C.c_int(0)
        check(_lib.load().shx_amr_run_to_end(self._h, batch, max_steps, C.byref(n)))
        if self.regrid_every > 0:
            self._refresh_tree()
        return n.value

    def _reconfigure(self, fn, cfg):
        cfg = replace(cfg or self.cfg, n=_level0(self.root, self.ndim), ndim=self.ndim)
        c = cfg.to_c()
        check(fn(self._h, C.byref(c)))
        self.cfg = cfg

    def write_params(self, cfg: SimConfig = None):
        """AmrSim.writeParams: the run-time fields of `cfg` (default self.cfg) from the next step on."""
        self._reconfigure(_lib.load().shx_amr_write_params, cfg)

    def build_hydro(self, cfg: SimConfig = None):
        """AmrSim.buildHydro: the block kernels of `cfg`'s limiter, Riemann solver and reconstruction."""
        self._reconfigure(_lib.load().shx_amr_build_hydro, cfg)

    def diagnostics(self) -> Diagnostics:
        out = np.zeros(24)
        check(_lib.load().shx_amr_diagnostics(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return diagnostics_from(out, self.clock())

    def flux_mismatch(self) -> dict:
        """AmrSim.fluxMismatch: same-level faces between blocks (and how many differ in any bit),
        coarse-fine faces, and the worst |coarse flux - mean of the fine fluxes|."""
        counts = (C.c_int64 * 3)()
        worst = C.c_double(0)
        check(_lib.load().shx_amr_flux_mismatch(self._h, counts, C.byref(worst)))
        return {"faces": counts[0], "mismatched": counts[1], "coarse_fine": counts[2], "worst_coarse_fine": worst.value}

    def profile(self, steps: int = 5) -> dict:
        ms = np.zeros(5)
        check(_lib.load().shx_amr_profile(self._h, steps, ms.ctypes.data_as(C.POINTER(C.c_double))))
        return {"assemble": ms[0], "clock": ms[1], "predict": ms[2], "update": ms[3], "step": ms[4]}

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().shx_amr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
