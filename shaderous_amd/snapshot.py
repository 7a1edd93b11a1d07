"""Snapshots and checkpoints of a run, and resuming from them: the reference's Snapshot
(src/io/snapshot.ts:18-60) and restart (SimConfig.restart, the clock from clockState,
src/sim.ts:1588-1593 and :2837-2838) for the hydro path.

- A ``Snapshot`` holds the gas as blocks of primitive variables (rho, press, vel1, vel2, vel3,
  r0; VARIABLES, snapshot.ts:9), each block as [z][y][x], with every block's refinement level and
  position at that level, the root grid and block sizes, the domain's edges, the time and the
  step count: the layout Athena++'s .athdf files store (``prim[variable][block][z][y][x]``,
  ``Levels``, ``LogicalLocations``; io/athdf.ts:4-12). A uniform grid is one block. A run
  restarted from one starts from the primitives, as the reference's does.
- A ``Checkpoint`` holds the conserved state as the GPU keeps it, with the time and step count: a
  run resumed from one continues bit for bit (the clock's dt comes from the state's rates; the
  Kahan term of the time restarts at zero, as clockState's does).

Files are .npz (numpy's zip of named arrays; read back without pickle). The .athdf container
itself needs an HDF5 library, which this image lacks: a writer maps ``Snapshot`` onto it field
for field.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

VARIABLES = ("rho", "press", "vel1", "vel2", "vel3", "r0")


@dataclass
class Snapshot:
    ndim: int
    root_size: tuple            # cells of the root grid per axis
    block_size: tuple           # cells of every block per axis
    lo: tuple                   # domain edges
    hi: tuple
    time: float
    steps: int
    levels: np.ndarray          # (blocks,) int32
    locations: np.ndarray       # (blocks, 3) int32, in blocks of the block's own level
    data: dict = field(default_factory=dict)   # variable -> (blocks, bz, by, bx) float32

    def save(self, path: str) -> None:
        meta = {"ndim": self.ndim, "root_size": list(self.root_size), "block_size": list(self.block_size),
                "lo": list(self.lo), "hi": list(self.hi), "time": self.time, "steps": self.steps}
        np.savez(path, meta=np.array(json.dumps(meta)), levels=self.levels, locations=self.locations,
                 **{"prim_" + k: v for k, v in self.data.items()})


@dataclass
class Checkpoint:
    u1: np.ndarray              # (cells, 4) as the GPU holds them
    u2: np.ndarray
    time: float
    steps: int
    blocks: np.ndarray | None = None   # the block tree, (blocks, 4) of (level, ix, iy, iz)
    root: tuple | None = None

    def save(self, path: str) -> None:
        arrays = {"u1": self.u1, "u2": self.u2, "clock": np.array([self.time, float(self.steps)])}
        if self.blocks is not None:
            arrays["blocks"] = self.blocks
            arrays["root"] = np.array(self.root, np.int32)
        np.savez(path, **arrays)


def load_snapshot(path: str) -> Snapshot:
    with np.load(path, allow_pickle=False) as f:
        meta = json.loads(str(f["meta"]))
        data = {k[5:]: f[k] for k in f.files if k.startswith("prim_")}
        return Snapshot(ndim=meta["ndim"], root_size=tuple(meta["root_size"]), block_size=tuple(meta["block_size"]),
                        lo=tuple(meta["lo"]), hi=tuple(meta["hi"]), time=meta["time"], steps=meta["steps"],
                        levels=f["levels"], locations=f["locations"], data=data)


def load_checkpoint(path: str) -> Checkpoint:
    with np.load(path, allow_pickle=False) as f:
        time, steps = f["clock"]
        blocks = f["blocks"] if "blocks" in f.files else None
        root = tuple(int(v) for v in f["root"]) if "root" in f.files else None
        return Checkpoint(u1=f["u1"], u2=f["u2"], time=float(time), steps=int(steps), blocks=blocks, root=root)


# ---- conserved <-> primitive (prim_from_cons, hydro/prim.wgsl, with the ideal gas) ----
def prims_of(u1, u2, gamma: float, p_floor: float = 0.0) -> dict:
    """Primitive variables of conserved cells (rows of 4): rho, press from the internal energy
    (with the pressure floor), velocity, and the scalar per unit mass."""
    u1 = np.asarray(u1, np.float64)
    u2 = np.asarray(u2, np.float64)
    rho = u1[:, 0]
    return {"rho": rho, "press": np.maximum((gamma - 1.0) * u2[:, 1], p_floor), "vel1": u1[:, 1] / rho,
            "vel2": u1[:, 2] / rho, "vel3": u1[:, 3] / rho, "r0": u2[:, 2] / rho}


def cons_of(prims: dict, gamma: float):
    """Conserved cells (u1, u2) of primitive variables: E = rho e + rho |v|^2 / 2, rho e = p / (gamma - 1)."""
    rho = np.asarray(prims["rho"], np.float64).ravel()
    v = [np.asarray(prims[k], np.float64).ravel() for k in ("vel1", "vel2", "vel3")]
    eint = np.asarray(prims["press"], np.float64).ravel() / (gamma - 1.0)
    r0 = np.asarray(prims.get("r0", np.zeros_like(rho)), np.float64).ravel()
    u1 = np.stack([rho, rho * v[0], rho * v[1], rho * v[2]], axis=1)
    u2 = np.stack([eint + 0.5 * rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2), eint, rho * r0, np.zeros_like(rho)], axis=1)
    return u1, u2


def _clock_of(sim):
    ck = sim.clock()
    # the clock holds the last step's dt beside the time it started from (Sim.diagnostics)
    return float(ck[1] + ck[0]), int(ck[3] + (1 if ck[0] > 0 else 0))


def _is_amr(sim) -> bool:
    return hasattr(sim, "blocks") and hasattr(sim, "root")


# ---- checkpoints ----
def checkpoint(sim) -> Checkpoint:
    """The run's conserved state, time and step count (a Sim or an AmrSim)."""
    u1, u2 = sim.read_state()
    time, steps = _clock_of(sim)
    if _is_amr(sim):
        return Checkpoint(u1=u1, u2=u2, time=time, steps=steps, blocks=np.array(sim.blocks), root=tuple(sim.root))
    return Checkpoint(u1=u1, u2=u2, time=time, steps=steps)


def resume(sim, ck: Checkpoint) -> None:
    """Put a checkpoint's state and clock on a run of the same grid (and tree)."""
    if _is_amr(sim) and (ck.blocks is None or not np.array_equal(np.asarray(ck.blocks), sim.blocks)):
        raise ValueError("the checkpoint's block tree is not the run's")
    sim.set_state(ck.u1, ck.u2, time=ck.time, steps=ck.steps)


# ---- snapshots ----
def capture(sim, box_lo: float | None = None) -> Snapshot:
    """The run's gas as a Snapshot: one block for a uniform grid, the leaves for an AmrSim."""
    cfg = sim.cfg
    u1, u2 = sim.read_state()
    time, steps = _clock_of(sim)
    n = tuple(int(v) for v in cfg.n)
    size = tuple(ni * cfg.dx for ni in n)
    lo = tuple(-s / 2 if box_lo is None else box_lo for s in size)
    hi = tuple(l + s for l, s in zip(lo, size))
    if _is_amr(sim):
        from .amr import BS, CELLS
        leaves = np.flatnonzero(sim.leaf_mask())
        sel = (leaves[:, None] * CELLS + np.arange(CELLS)[None, :]).ravel()
        prims = prims_of(u1[sel], u2[sel], cfg.gamma, cfg.p_floor)
        data = {k: prims[k].astype(np.float32).reshape(len(leaves), BS, BS, BS) for k in VARIABLES}
        blocks = sim.blocks[leaves]
        return Snapshot(ndim=3, root_size=n, block_size=(BS, BS, BS), lo=lo, hi=hi, time=time, steps=steps,
                        levels=blocks[:, 0].astype(np.int32), locations=blocks[:, 1:4].astype(np.int32), data=data)
    prims = prims_of(u1, u2, cfg.gamma, cfg.p_floor)
    data = {k: prims[k].astype(np.float32).reshape(1, n[2], n[1], n[0]) for k in VARIABLES}
    return Snapshot(ndim=cfg.ndim, root_size=n, block_size=n, lo=lo, hi=hi, time=time, steps=steps,
                    levels=np.zeros(1, np.int32), locations=np.zeros((1, 3), np.int32), data=data)


def restart(sim, snap: Snapshot) -> None:
    """Start a run from a Snapshot (SimConfig.restart): its primitives as conserved state, its
    time and step count on the clock. An AmrSim's tree must hold the snapshot's blocks as
    leaves (in any order); its parents' cells are left zero."""
    gamma = sim.cfg.gamma
    if _is_amr(sim):
        from .amr import CELLS
        u1 = np.zeros((sim.ncell, 4))
        u2 = np.zeros((sim.ncell, 4))
        slot = {tuple(b): i for i, b in enumerate(sim.blocks.tolist())}
        leaf = sim.leaf_mask()
        for j in range(len(snap.levels)):
            key = (int(snap.levels[j]), *(int(v) for v in snap.locations[j]))
            i = slot.get(key)
            if i is None or not leaf[i]:
                raise ValueError(f"snapshot block {key} is not a leaf of the run's tree")
            c1, c2 = cons_of({k: snap.data[k][j] for k in VARIABLES}, gamma)
            u1[i * CELLS:(i + 1) * CELLS] = c1
            u2[i * CELLS:(i + 1) * CELLS] = c2
    else:
        if len(snap.levels) != 1 or tuple(snap.block_size) != tuple(int(v) for v in sim.cfg.n):
            raise ValueError("a uniform run restarts from a one-block snapshot of its own size")
        u1, u2 = cons_of({k: snap.data[k][0] for k in VARIABLES}, gamma)
    sim.set_state(u1, u2, time=snap.time, steps=snap.steps)
