"""The batched ensemble of independent 2D grids (shx_ens_*), on an MI355X.

- every instance gives the same bits as a standalone uniform-grid run of the plain path with the
  same configuration (its own dt: the instances' clocks are independent);
- small instances against the FP32 oracle, per step, within the FP32 bar (L1 <= 1e-5);
- the ensemble workload (64 seeded 2D four-quadrant Riemann problems, sizes drawn from
  {64, 128, 256, 512}^2, gamma 1.4, outflow, 200 steps): FP32 against the FP64 twin of the same
  kernels for every instance, and against the FP64 oracle for the 64^2 instances.
"""
import json
import os

import numpy as np
import pytest

from shaderous_amd import problems
from shaderous_amd.ensemble import Ensemble
from shaderous_amd.sim import Sim, SimConfig, config_for

oracle = pytest.importorskip("oracle")
from oracle import params, wgsl32, wgsl64  # noqa: E402

pytestmark = pytest.mark.gpu

OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")


def shared_cfg(precision=32, limiter="mc", isentropic=False, riemann="hllc"):
    return SimConfig(n=(64, 64, 1), ndim=2, dx=1.0 / 64, gamma=1.4, bc_lo=(1, 1, 0), bc_hi=(1, 1, 0), limiter=limiter,
                     isentropic=isentropic, cfl=0.4, precision=precision, riemann=riemann)


def l1(a, b):
    a1, a2 = (np.asarray(x, np.float64) for x in a)
    b1, b2 = (np.asarray(x, np.float64) for x in b)
    return {"rho": float(np.abs(a1[:, 0] - b1[:, 0]).sum() / np.abs(b1[:, 0]).sum()),
            "mom": float(np.abs(a1[:, 1:3] - b1[:, 1:3]).sum() / max(np.abs(b1[:, 1:3]).sum(), 1e-300)),
            "E": float(np.abs(a2[:, 0] - b2[:, 0]).sum() / np.abs(b2[:, 0]).sum())}


# the ensemble's work shapes, by cells: k_ens_fused's tiles (threads 32 x 16, 32 x 32, 20 x 20)
# and k_ens_rows' strips (60 cells x 8, 16 or 32 rows)
ENS_TILES = [(28, 12), (28, 28), (16, 16), (60, 8), (60, 16), (60, 32)]


@pytest.mark.parametrize("lim,isen,tile", [("mc", False, t) for t in ENS_TILES] + [("koren", True, t) for t in ENS_TILES])
def test_ensemble_equals_standalone_runs(lim, isen, tile):
    equals_standalone(lim, isen, tile, "hllc")


# every other limiter and solver of the reference (src/problems/types.ts:81-84), on the row march
# and on the tile kernel
@pytest.mark.parametrize("tile", [(60, 16), (16, 16)])
@pytest.mark.parametrize("lim,rs,isen", [("vanleer", "hllc", False), ("minmod", "hll", False), ("koren", "hll", True),
                                         ("mc", "hll", False)])
def test_ensemble_other_schemes_equal_standalone_runs(lim, rs, isen, tile):
    equals_standalone(lim, isen, tile, rs)


def equals_standalone(lim, isen, tile, rs):
    """Every instance of an ensemble gives the bits of its own plain-path Sim."""
    sizes = [(64, 64), (100, 37), (128, 128), (30, 90), (200, 64)]
    cfg = shared_cfg(limiter=lim, isentropic=isen, riemann=rs)
    cfg.tile = (tile[0], tile[1], 0)
    ens = Ensemble(cfg, sizes)
    pbs = []
    try:
        for i, (nx, ny) in enumerate(sizes):
            pb = problems.riemann_quadrants(nx, seed=100 + i, limiter=lim, isentropic=isen)
            if ny != nx:
                x, y, _ = problems.centers((nx, ny, 1), 1.0)
                st = problems.quadrant_states(100 + i)
                q = (x > 0).astype(int) + 2 * (y > 0).astype(int)
                pb = problems.Problem(id="q", n=(nx, ny, 1), box=1.0, gamma=1.4, bc_lo=(1, 1, 0), bc_hi=(1, 1, 0),
                                      cfl=0.4, limiter=lim, isentropic=isen, riemann=rs,
                                      prims={"rho": st[q, 0], "vx": st[q, 1], "vy": st[q, 2], "p": st[q, 3]})
            pb.riemann = rs
            pbs.append(pb)
            ens.set_state(i, *pb.state())
        ens.step(7)
        got = [ens.read_state(i) for i in range(len(sizes))]
        clocks = [ens.clock(i) for i in range(len(sizes))]
    finally:
        ens.destroy()
    dts = set()
    for i, pb in enumerate(pbs):
        sim = Sim(config_for(pb, path="plain"))
        try:
            sim.set_state(*pb.state())
            sim.step(7)
            ref = sim.read_state()
            ck = sim.clock()
        finally:
            sim.destroy()
        assert clocks[i][1] == ck[1] and clocks[i][0] == ck[0], (i, clocks[i][:2], ck[:2])
        dts.add(float(ck[0]))
        same = np.array_equal(got[i][0], ref[0]) and np.array_equal(got[i][1], ref[1])
        assert same, (i, l1(got[i], ref))
    assert len(dts) > 1   # the instances took their own steps


def test_ensemble_matches_oracle_per_step():
    sizes = [(48, 48), (40, 72)]
    cfg = shared_cfg()
    ens = Ensemble(cfg, sizes)
    states = []
    try:
        for i, (nx, ny) in enumerate(sizes):
            x, y, _ = problems.centers((nx, ny, 1), 1.0)
            st = problems.quadrant_states(7 + i)
            q = (x > 0).astype(int) + 2 * (y > 0).astype(int)
            u1, u2 = problems.cons_from_prims({"rho": st[q, 0], "vx": st[q, 1], "vy": st[q, 2], "p": st[q, 3]}, 1.4)
            ens.set_state(i, u1, u2)
            states.append((u1.copy(), u2.copy(), np.zeros(4, np.float32)))
        for _ in range(5):
            ens.step(1)
            for i, (nx, ny) in enumerate(sizes):
                prm = params((nx, ny, 1), ndim=2, dx=1.0 / nx, gamma=1.4, cfl=0.4, limiter="mc", bc_lo=(1, 1, 0),
                             bc_hi=(1, 1, 0))
                o1, o2, clk = states[i]
                wgsl32.step(prm, o1, o2, clk)
                d = l1(ens.read_state(i), (o1, o2))
                assert max(d.values()) <= 1e-5, (i, d)
                assert abs(ens.clock(i)[0] - clk[0]) <= 1e-5 * clk[0]
    finally:
        ens.destroy()


def test_ensemble_workload_precision():
    """64 seeded quadrant problems, sizes from {64, 128, 256, 512}^2, 200 steps: FP32 vs the FP64
    twin per instance; FP32 vs the FP64 oracle for the 64^2 instances."""
    ns = problems.ensemble_sizes(64, seed=42)
    sizes = [(n, n) for n in ns]
    pbs = [problems.riemann_quadrants(n, seed=1000 + i) for i, n in enumerate(ns)]
    out = {}
    for precision in (32, 64):
        ens = Ensemble(shared_cfg(precision), sizes)
        try:
            dtype = np.float64 if precision == 64 else np.float32
            for i, pb in enumerate(pbs):
                ens.set_state(i, *pb.state(dtype))
            ens.step(200)
            out[precision] = [ens.read_state(i) for i in range(len(sizes))]
        finally:
            ens.destroy()
    per = [l1(out[32][i], out[64][i]) for i in range(len(sizes))]
    worst = {k: max(p[k] for p in per) for k in ("rho", "mom", "E")}
    orc = {}
    for i, n in enumerate(ns):
        if n != 64:
            continue
        prm = params((64, 64, 1), ndim=2, dx=1.0 / 64, gamma=1.4, cfl=0.4, limiter="mc", bc_lo=(1, 1, 0),
                     bc_hi=(1, 1, 0))
        u1, u2 = pbs[i].state(np.float64)
        wgsl64.run(prm, u1, u2, 200)
        orc[i] = l1(out[32][i], (u1, u2))
    data = {"sizes": ns, "l1_fp32_vs_fp64_gpu": per, "worst": worst, "l1_fp32_vs_fp64_oracle_64^2": orc}
    print("ensemble worst", worst, "oracle", orc)
    if os.path.isdir(OUT):
        json.dump(data, open(os.path.join(OUT, "ensemble_precision.json"), "w"), indent=1)
    assert max(worst.values()) <= 1e-5, worst
    assert orc and max(max(v.values()) for v in orc.values()) <= 1e-5
