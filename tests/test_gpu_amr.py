"""Parity of the HIP block-AMR path (through the C ABI) with the block-AMR oracle, on an MI355X.

Bar: per-step L1 relative difference of leaves' density, momentum and energy <= 1e-5 in FP32
against the FP32 oracle (FP64 against FP64 to 1e-12); conserved totals to round-off; the
per-level and batched launch forms equal bit for bit; the reference's AMR validation criteria
(src/validation/tests.ts:1202-1263): blocks without refinement within 1e-5 of the uniform grid,
and a refined Sod tube within L1 0.003 of the exact solution with drift below 1e-6.
"""
import math

import numpy as np
import pytest

from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import Sim, SimConfig

oracle = pytest.importorskip("oracle")
from oracle import AmrTable, exact_riemann, params, wgsl32, wgsl64  # noqa: E402

pytestmark = pytest.mark.gpu

BS = 16
L1_TOL = 1e-5


def blob(x, y, z):
    r2 = x * x + y * y + z * z
    g = np.exp(-r2 / 0.02)
    return {"rho": 1.0 + 0.8 * g, "vx": 0.3 + 0.2 * np.sin(2 * np.pi * y) - 2.0 * x * g,
            "vy": -0.2 * x - 2.0 * y * g, "vz": 0.1 * np.cos(2 * np.pi * z) - 2.0 * z * g,
            "p": 1.0 + 2.0 * np.exp(-r2 / 0.01), "s": (x > 0).astype(float)}


def cfg_for(root, precision=32, limiter="mc", isentropic=False, bc_lo=(0, 0, 0), bc_hi=(0, 0, 0), cfl=0.4,
            gamma=1.4, t_end=math.inf):
    n = tuple(BS * r for r in root)
    return SimConfig(n=n, ndim=3, dx=1.0 / n[0], gamma=gamma, bc_lo=bc_lo, bc_hi=bc_hi, limiter=limiter,
                     isentropic=isentropic, cfl=cfl, precision=precision, t_end=t_end)


def orc_prm(cfg):
    return params(cfg.n, ndim=3, dx=cfg.dx, gamma=cfg.gamma, cfl=cfg.cfl, bc_lo=cfg.bc_lo, bc_hi=cfg.bc_hi,
                  limiter=cfg.limiter, riemann=cfg.riemann, isentropic=cfg.isentropic,
                  t_end=cfg.t_end if math.isfinite(cfg.t_end) else 1e30)


def refined_tree():
    # root 2^3, levels 0..2 around one corner (balanced by the builder)
    return amr_tree.build_tree((2, 2, 2), 2, lambda L, c: c == (0, 0, 0) or (L == 1 and c == (1, 1, 1)))


def small_blast():
    return amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))[0]


def leaf_rows(blocks):
    return amr_tree.leaf_mask(blocks).repeat(BS ** 3)


def l1rel(a, b, sel):
    out = {}
    out["rho"] = np.abs(a[0][sel, 0] - b[0][sel, 0]).sum() / np.abs(b[0][sel, 0]).sum()
    out["mom"] = np.abs(a[0][sel, 1:4] - b[0][sel, 1:4]).sum() / max(np.abs(b[0][sel, 1:4]).sum(), 1e-300)
    out["E"] = np.abs(a[1][sel, 0] - b[1][sel, 0]).sum() / np.abs(b[1][sel, 0]).sum()
    return out


def run_pair(blocks, root, cfg, steps, path="passes", compensated=True, ic=blob):
    """Step the GPU and the oracle of the same precision side by side; worst per-step L1."""
    impl = wgsl64 if cfg.precision == 64 else wgsl32
    dx0 = cfg.dx
    u1, u2 = amr_tree.blocks_state(blocks, dx0, ic, cfg.gamma, impl.dtype)
    o1, o2 = u1.copy(), u2.copy()
    t = AmrTable(blocks, root, cfg.bc_lo, cfg.bc_hi)
    prm = orc_prm(cfg)
    clk = np.zeros(4, impl.dtype)
    sel = leaf_rows(blocks)
    worst = {"rho": 0.0, "mom": 0.0, "E": 0.0}
    sim = AmrSim(cfg, blocks, root, path=path, compensated=compensated)
    try:
        assert np.array_equal(sim.table(), t.table())
        sim.set_state(u1, u2)
        for _ in range(steps):
            sim.step(1)
            impl.amr_step(prm, t, dx0, o1, o2, clk, compensated)
            g = sim.read_state()
            d = l1rel(g, (o1, o2), sel)
            for k in worst:
                worst[k] = max(worst[k], d[k])
            gc = sim.clock()
            assert abs(gc[0] - clk[0]) <= (1e-12 if cfg.precision == 64 else 1e-5) * clk[0], (gc[0], clk[0])
        state = sim.read_state()
    finally:
        sim.destroy()
    return worst, state, (o1, o2)


def test_amr_table_matches_oracle_blast():
    blocks = small_blast()
    cfg = cfg_for((4, 4, 4), bc_lo=(2, 1, 0), bc_hi=(1, 2, 0))
    sim = AmrSim(cfg, blocks, (4, 4, 4))
    try:
        assert np.array_equal(sim.table(), AmrTable(blocks, (4, 4, 4), (2, 1, 0), (1, 2, 0)).table())
        st = sim.stats()
        assert st["blocks"] == len(blocks) and st["leaves"] == int(amr_tree.leaf_mask(blocks).sum())
        assert st["levels"] == 3 and st["match_leaves"] > 0
    finally:
        sim.destroy()


@pytest.mark.parametrize("path", ["fused", "passes"])
@pytest.mark.parametrize("compensated", [True, False])
@pytest.mark.parametrize("lim,isen", [("mc", False), ("koren", True)])
def test_amr_fp32_matches_oracle_refined(lim, isen, compensated, path):
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2), limiter=lim, isentropic=isen)
    worst, _, _ = run_pair(blocks, (2, 2, 2), cfg, 4, path=path, compensated=compensated)
    print("L1", worst)
    assert max(worst.values()) <= L1_TOL, worst


@pytest.mark.parametrize("path", ["fused", "passes"])
@pytest.mark.parametrize("lim,rs,isen", [("vanleer", "hllc", False), ("minmod", "hll", False), ("koren", "hll", True),
                                         ("mc", "hll", False)])
def test_amr_other_schemes_match_oracle(lim, rs, isen, path):
    """Every limiter and solver the reference offers (src/problems/types.ts:81-84) on blocks."""
    from dataclasses import replace
    blocks = refined_tree()
    cfg = replace(cfg_for((2, 2, 2), limiter=lim, isentropic=isen), riemann=rs)
    worst, _, _ = run_pair(blocks, (2, 2, 2), cfg, 3, path=path)
    print("L1", lim, rs, isen, path, worst)
    assert max(worst.values()) <= L1_TOL, worst


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_amr_fp32_matches_oracle_blast_walls(path):
    blocks = small_blast()
    cfg = cfg_for((4, 4, 4), bc_lo=(2, 1, 0), bc_hi=(1, 2, 0))

    def ic(x, y, z):
        r = np.sqrt((x - 0.03) ** 2 + (y + 0.02) ** 2 + (z - 0.05) ** 2)
        return {"rho": np.ones_like(x), "p": np.where(r < 0.2, 10.0, 1.0), "vx": 0.3 * np.sin(6 * y)}
    worst, _, _ = run_pair(blocks, (4, 4, 4), cfg, 3, path=path, ic=ic)
    print("L1", worst)
    assert max(worst.values()) <= L1_TOL, worst


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_amr_fp64_matches_oracle(path):
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2), precision=64)
    worst, _, _ = run_pair(blocks, (2, 2, 2), cfg, 3, path=path)
    print("L1", worst)
    assert max(worst.values()) <= 1e-12, worst


def test_amr_paths_equal_bitwise():
    """The fused form, the batched passes and the per-level passes give the same bits."""
    blocks = small_blast()
    cfg = cfg_for((4, 4, 4))
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    out = {}
    for path in ("passes", "levels", "fused"):
        sim = AmrSim(cfg, blocks, (4, 4, 4), path=path)
        try:
            sim.set_state(u1, u2)
            sim.step(5)
            out[path] = sim.read_state()
            if path == "levels":
                assert sim.stats()["launches_per_step"] > 5
        finally:
            sim.destroy()
    sel = leaf_rows(blocks)
    assert np.array_equal(out["passes"][0][sel], out["levels"][0][sel])
    assert np.array_equal(out["passes"][1][sel], out["levels"][1][sel])
    d = l1rel(out["fused"], out["passes"], sel)
    print("fused vs passes", d, "bitwise", np.array_equal(out["fused"][0][sel], out["passes"][0][sel]))
    assert np.array_equal(out["fused"][0][sel], out["passes"][0][sel])
    assert np.array_equal(out["fused"][1][sel], out["passes"][1][sel])


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_amr_mixed_boundaries_match_oracle(path):
    """Every boundary kind on one refined grid: x outflow-only below and outflow above, y periodic
    below and outflow above (periodic on one face only), z reflecting below and outflow-only
    above; against the FP32 oracle per step, and the fused and pass forms bit for bit."""
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2), bc_lo=(3, 0, 2), bc_hi=(1, 1, 3))

    def ic(x, y, z):
        r2 = x * x + y * y + z * z
        g = np.exp(-r2 / 0.02)
        # flow into the domain through the outflow-only faces, so their clamps act
        return {"rho": 1.0 + 0.8 * g, "vx": 0.6 - 2.0 * x * g, "vy": 0.3 * np.sin(2 * np.pi * x),
                "vz": -0.5 - 2.0 * z * g, "p": 1.0 + 2.0 * np.exp(-r2 / 0.01)}
    worst, state, _ = run_pair(blocks, (2, 2, 2), cfg, 4, path=path, ic=ic)
    print("L1", path, worst)
    assert max(worst.values()) <= L1_TOL, worst
    if path == "fused":
        other = run_pair(blocks, (2, 2, 2), cfg, 4, path="passes", ic=ic)[1]
        sel = leaf_rows(blocks)
        assert np.array_equal(state[0][sel], other[0][sel]) and np.array_equal(state[1][sel], other[1][sel])


@pytest.mark.parametrize("precision,tol", [(32, 2e-6), (64, 1e-13)])
def test_amr_conservation(precision, tol):
    blocks = small_blast()
    cfg = cfg_for((4, 4, 4), precision=precision)
    dtype = np.float64 if precision == 64 else np.float32
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma, dtype)
    sim = AmrSim(cfg, blocks, (4, 4, 4))
    try:
        sim.set_state(u1, u2)
        d0 = sim.diagnostics()
        sim.step(20)
        d1 = sim.diagnostics()
    finally:
        sim.destroy()
    assert abs(d1.mass - d0.mass) / d0.mass < tol
    assert abs(d1.e_tot - d0.e_tot) / d0.e_tot < tol
    for a in range(3):
        assert abs(d1.momentum[a] - d0.momentum[a]) < tol * d0.mass


def test_amr_unrefined_matches_uniform():
    """amr-unrefined (src/validation/tests.ts:1202-1229): a block grid with no refinement against
    the uniform grid, density in every cell within 1e-5 relative, the same steps."""
    root = (2, 2, 2)
    cfg = cfg_for(root)
    blocks = np.asarray([(0, x, y, z) for z in range(2) for y in range(2) for x in range(2)], np.int32)
    n = cfg.n
    u1, u2 = problems.cons_from_prims(blob(*problems.centers(n, 1.0)), cfg.gamma)
    b1, b2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    uni = Sim(SimConfig(**{**cfg.__dict__, "path": "plain"}))
    amr = AmrSim(cfg, blocks, root)
    try:
        uni.set_state(u1, u2)
        amr.set_state(b1, b2)
        uni.step(10)
        amr.step(10)
        g1, _ = uni.read_state()
        a1, _ = amr.read_state()
        assert abs(uni.clock()[1] - amr.clock()[1]) <= 1e-6 * uni.clock()[1]
    finally:
        uni.destroy()
        amr.destroy()
    g = g1.reshape(32, 32, 32, 4)
    for i, (_, bx, by, bz) in enumerate(blocks):
        ref = g[16 * bz:16 * bz + 16, 16 * by:16 * by + 16, 16 * bx:16 * bx + 16, 0].reshape(-1)
        got = a1[i * 4096:(i + 1) * 4096, 0]
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-5


def test_amr_sod_tube():
    """amr-shock-tube (src/validation/tests.ts:1230-1263) on a static mesh: Sod along x in 3D,
    128 level-0 cells refined to 256 over the region the waves cross by t = 0.2; L1 density error
    < 0.003, mass and energy drift < 1e-6, fewer cells than the uniform 256 grid."""
    root = (8, 1, 1)
    cfg = cfg_for(root, bc_lo=(1, 0, 0), bc_hi=(1, 0, 0), cfl=0.8, t_end=0.2, limiter="koren", isentropic=True)
    # level-0 blocks 2..5 cover x in [-0.25, 0.25); 6 covers [0.25, 0.375): refine 2..6
    blocks = amr_tree.build_tree(root, 1, lambda L, c: 2 <= c[0] <= 6, periodic=(False, True, True))
    left, right = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)

    def ic(x, y, z):
        L = x < 0
        return {"rho": np.where(L, left[0], right[0]), "p": np.where(L, left[2], right[2])}
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, ic, 1.4)
    sim = AmrSim(cfg, blocks, root)
    try:
        sim.set_state(u1, u2)
        d0 = sim.diagnostics()
        sim.run_to(0.2, max_steps=2000)
        d1 = sim.diagnostics()
        c = sim.read_leaves()
    finally:
        sim.destroy()
    assert abs(d1.t - 0.2) < 1e-6
    ex = np.asarray([exact_riemann(left, right, 1.4, x / 0.2)[0] for x in c["x"]])
    vol = c["dx"] ** 3
    l1 = float(np.sum(np.abs(c["u1"][:, 0] - ex) * vol) / vol.sum())
    print("sod L1", l1, "cells", c["count"])
    assert l1 < 0.003
    assert abs(d1.mass - d0.mass) / d0.mass < 1e-6
    assert abs(d1.e_tot - d0.e_tot) / d0.e_tot < 1e-6
    assert c["count"] < 256 * 32 * 32


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_amr_flux_mismatch(path):
    """AmrSim.fluxMismatch (amr-shock-tube's criterion, src/validation/tests.ts:1230-1263): every
    same-level face between blocks bit-identical from both sides, coarse-fine faces within 1e-6
    of the mean of the fine fluxes; the state and clock are left as they were."""
    blocks = small_blast()
    cfg = cfg_for((4, 4, 4), bc_lo=(2, 1, 0), bc_hi=(1, 2, 0))
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    sim = AmrSim(cfg, blocks, (4, 4, 4), path=path)
    try:
        sim.set_state(u1, u2)
        sim.step(3)
        before = sim.read_state()
        ck = sim.clock()
        fm = sim.flux_mismatch()
        after = sim.read_state()
        assert np.array_equal(before[0], after[0]) and np.array_equal(sim.clock(), ck)
        sim.step(1)
    finally:
        sim.destroy()
    print("flux", fm)
    assert fm["faces"] > 0 and fm["mismatched"] == 0
    assert fm["coarse_fine"] > 0 and fm["worst_coarse_fine"] < 1e-6


def test_amr_errors():
    cfg = cfg_for((2, 2, 2))
    bad = [(0, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    bad += [(1, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    bad += [(2, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    from shaderous_amd._lib import ShxError
    with pytest.raises(ShxError, match="2:1"):
        AmrSim(cfg, np.asarray(bad), (2, 2, 2))
    with pytest.raises(ShxError, match="children"):
        AmrSim(cfg, np.asarray(bad[:8] + bad[8:12]), (2, 2, 2))
    with pytest.raises(ShxError, match="root"):
        AmrSim(cfg, np.asarray(bad[:7]), (2, 2, 2))
