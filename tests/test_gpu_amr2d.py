"""Parity of the HIP 2D block-AMR path (quadtrees of 16^2-cell blocks, shx_amr with ndim 2)
through the C ABI with the 2D block-AMR oracle (tests/test_oracle_amr2d.py pins it), on an MI355X.

Bar, as for 3D: per-step L1 relative difference of leaves' density, momentum and energy <= 1e-5
in FP32 against the FP32 oracle (FP64 against FP64 to 1e-12); conserved totals to round-off; the
per-level and batched forms equal bit for bit; an unrefined quadtree within 1e-5 of the uniform
2D grid (src/validation/tests.ts:1202-1229); same-level faces bit-identical from both sides and
coarse-fine faces within 1e-6 of the mean of the fine fluxes; diagnostics equal to host FP64 sums
of the state; re-meshing and per-block timesteps as the oracle's.
"""
import math

import numpy as np
import pytest

from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import Sim, SimConfig

oracle = pytest.importorskip("oracle")
from oracle import AmrTable, params, wgsl32, wgsl64  # noqa: E402

pytestmark = pytest.mark.gpu

BS = 16
ROOT = (4, 4, 1)
DX0 = 1.0 / 64


def flow(x, y, z):
    """Density and pressure blobs (the pressure's a blast) around (-1/8, -1/8), in a flow that
    expands with a clear divergence: the shock detector's velocity convergence at the rim ghosts
    beside a coarser block (where an edge texel is a neighbour) keeps a definite sign, which the
    GPU's v = m * (1 / rho) and the oracle's m / rho would otherwise round differently."""
    r2 = (x + 0.125) ** 2 + (y + 0.125) ** 2
    return {"rho": 1.0 + 0.6 * np.exp(-r2 / 0.01), "vx": 0.4 + 0.3 * x, "vy": 0.3 + 0.2 * y, "vz": 0.0 * x,
            "p": 0.2 + 3.0 * np.exp(-r2 / 0.003), "s": (x > -0.1).astype(float)}


def quadtree():
    return amr_tree.build_tree(ROOT, 2, lambda L, c: (L == 0 and c == (1, 1, 0)) or (L == 1 and c == (3, 3, 0)),
                               ndim=2)


def cfg2(precision=32, limiter="mc", isentropic=False, riemann="hllc", bc_lo=(0, 0, 0), bc_hi=(0, 0, 0),
         t_end=math.inf):
    return SimConfig(n=(64, 64, 1), ndim=2, dx=DX0, gamma=1.4, bc_lo=bc_lo, bc_hi=bc_hi, limiter=limiter,
                     riemann=riemann, isentropic=isentropic, cfl=0.4, precision=precision, t_end=t_end)


def orc_prm(cfg):
    return params((64, 64, 1), ndim=2, dx=cfg.dx, gamma=cfg.gamma, cfl=cfg.cfl, bc_lo=cfg.bc_lo, bc_hi=cfg.bc_hi,
                  limiter=cfg.limiter, riemann=cfg.riemann, isentropic=cfg.isentropic,
                  t_end=cfg.t_end if math.isfinite(cfg.t_end) else 1e30)


def leaf_rows(blocks):
    return amr_tree.leaf_mask(blocks).repeat(BS ** 2)


def l1rel(a, b, sel):
    out = {}
    out["rho"] = np.abs(a[0][sel, 0] - b[0][sel, 0]).sum() / np.abs(b[0][sel, 0]).sum()
    out["mom"] = np.abs(a[0][sel, 1:4] - b[0][sel, 1:4]).sum() / max(np.abs(b[0][sel, 1:4]).sum(), 1e-300)
    out["E"] = np.abs(a[1][sel, 0] - b[1][sel, 0]).sum() / np.abs(b[1][sel, 0]).sum()
    return out


def run_pair(blocks, cfg, steps, path="passes", compensated=True):
    impl = wgsl64 if cfg.precision == 64 else wgsl32
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, cfg.gamma, impl.dtype, ndim=2)
    o1, o2 = u1.copy(), u2.copy()
    t = AmrTable(blocks, ROOT, cfg.bc_lo, cfg.bc_hi, ndim=2)
    prm = orc_prm(cfg)
    clk = np.zeros(4, impl.dtype)
    sel = leaf_rows(blocks)
    worst = {"rho": 0.0, "mom": 0.0, "E": 0.0}
    sim = AmrSim(cfg, blocks, ROOT, path=path, compensated=compensated)
    try:
        assert np.array_equal(sim.table(), t.table())
        sim.set_state(u1, u2)
        for _ in range(steps):
            sim.step(1)
            impl.amr_step(prm, t, DX0, o1, o2, clk, compensated)
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


@pytest.mark.parametrize("path", ["passes", "levels"])
@pytest.mark.parametrize("compensated", [True, False])
@pytest.mark.parametrize("lim,isen", [("mc", False), ("koren", True)])
def test_2d_fp32_matches_oracle(lim, isen, compensated, path):
    worst, _, _ = run_pair(quadtree(), cfg2(limiter=lim, isentropic=isen), 6, path, compensated)
    print("2d fp32", lim, path, compensated, worst)
    assert max(worst.values()) <= 1e-5, worst


@pytest.mark.parametrize("lim,rs,isen", [("vanleer", "hllc", False), ("minmod", "hll", False), ("koren", "hll", True),
                                         ("koren", "hllc", False), ("mc", "hll", False)])
def test_2d_other_schemes_match_oracle(lim, rs, isen):
    worst, _, _ = run_pair(quadtree(), cfg2(limiter=lim, riemann=rs, isentropic=isen), 4)
    assert max(worst.values()) <= 1e-5, (lim, rs, worst)


@pytest.mark.parametrize("lim,isen", [("mc", False), ("koren", True)])
def test_2d_fp64_matches_oracle(lim, isen):
    worst, _, _ = run_pair(quadtree(), cfg2(precision=64, limiter=lim, isentropic=isen), 6)
    assert max(worst.values()) <= 1e-12, worst


def test_2d_walls_match_oracle():
    worst, _, _ = run_pair(quadtree(), cfg2(bc_lo=(1, 2, 0), bc_hi=(2, 1, 0)), 6)
    assert max(worst.values()) <= 1e-5, worst


def test_2d_paths_equal_bitwise():
    blocks = quadtree()
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, np.float32, ndim=2)
    out = {}
    for path in ("passes", "levels"):
        sim = AmrSim(cfg2(), blocks, ROOT, path=path)
        try:
            sim.set_state(u1, u2)
            sim.step(8)
            out[path] = sim.read_state()
            st = sim.stats()
        finally:
            sim.destroy()
        assert st["levels"] == 3
    assert np.array_equal(out["passes"][0], out["levels"][0]) and np.array_equal(out["passes"][1], out["levels"][1])


def totals(blocks, u1, u2):
    leaf = amr_tree.leaf_mask(blocks)
    vol = np.repeat((DX0 / (1 << blocks[:, 0].astype(np.int64))) ** 2 * leaf, BS ** 2)
    return np.concatenate([(u1.astype(np.float64) * vol[:, None]).sum(0), [(u2[:, 0].astype(np.float64) * vol).sum()]])


@pytest.mark.parametrize("precision,tol", [(32, 2e-6), (64, 1e-13)])
def test_2d_conservation(precision, tol):
    blocks = quadtree()
    dtype = np.float64 if precision == 64 else np.float32
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, dtype, ndim=2)
    c0 = totals(blocks, u1, u2)
    sim = AmrSim(cfg2(precision=precision), blocks, ROOT, path="passes")
    try:
        sim.set_state(u1, u2)
        sim.step(20)
        g1, g2 = sim.read_state()
        d = sim.diagnostics()
    finally:
        sim.destroy()
    c1 = totals(blocks, g1, g2)
    scale = np.abs(c0) + np.array([0, 1, 1, 1, 0]) * c0[0]
    assert np.all(np.abs(c1 - c0) <= tol * scale), (c1 - c0) / scale
    # the diagnostics' FP64 sums over leaves (cells of area h^2) against the host's
    assert abs(d.mass - c1[0]) <= 1e-12 * c1[0]
    assert abs(d.e_tot - c1[4]) <= 1e-12 * abs(c1[4])


def test_2d_unrefined_matches_uniform():
    """amr-unrefined in 2D: a quadtree with no refinement against the uniform 2D grid, density in
    every cell within 1e-5 relative after the same 10 steps."""
    cfg = cfg2()
    roots = np.asarray([(0, i, j, 0) for j in range(4) for i in range(4)], np.int32)
    u1, u2 = problems.cons_from_prims(flow(*problems.centers((64, 64, 1), 1.0)), 1.4)
    b1, b2 = amr_tree.blocks_state(roots, DX0, flow, 1.4, ndim=2)
    uni = Sim(SimConfig(**{**cfg.__dict__, "path": "plain"}))
    amr = AmrSim(cfg, roots, ROOT, path="passes")
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
    g = g1.reshape(64, 64, 4)
    for i, (_, bx, by, _) in enumerate(roots):
        ref = g[16 * by:16 * by + 16, 16 * bx:16 * bx + 16, 0].reshape(-1)
        got = a1[i * 256:(i + 1) * 256, 0]
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-5


def test_2d_flux_mismatch():
    blocks = quadtree()
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, ndim=2)
    sim = AmrSim(cfg2(bc_lo=(2, 1, 0), bc_hi=(1, 2, 0)), blocks, ROOT, path="passes")
    try:
        sim.set_state(u1, u2)
        sim.step(3)
        before = sim.read_state()
        fm = sim.flux_mismatch()
        after = sim.read_state()
        assert np.array_equal(before[0], after[0])
    finally:
        sim.destroy()
    print("flux 2d", fm)
    assert fm["faces"] > 0 and fm["mismatched"] == 0
    assert fm["coarse_fine"] > 0 and fm["worst_coarse_fine"] < 1e-6


@pytest.mark.parametrize("precision,tol", [(32, 1e-6), (64, 1e-13)])
def test_2d_regrid_matches_oracle(precision, tol):
    """Re-meshing on the GPU (split every leaf below level 2, then merge back) against the oracle's
    regrid of the same state and maps."""
    blocks = quadtree()
    impl = wgsl64 if precision == 64 else wgsl32
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, impl.dtype, ndim=2)
    leaf = amr_tree.leaf_mask(blocks)
    up = amr_tree.adapt_tree(blocks, ROOT, np.where(leaf, 1.0, np.nan), 0.5, -1.0, 2, ndim=2)
    k, s, o = amr_tree.regrid_map(blocks, up)
    w1, w2 = impl.amr_regrid(orc_prm(cfg2(precision)), AmrTable(blocks, ROOT, ndim=2), u1, u2, k, s, o)
    sim = AmrSim(cfg2(precision), blocks, ROOT, path="passes")
    try:
        sim.set_state(u1, u2)
        sim.regrid(up)
        g1, g2 = sim.read_state()
        assert np.array_equal(sim.table(), AmrTable(up, ROOT, ndim=2).table())
        sim.regrid(blocks)
        b1, _ = sim.read_state()
    finally:
        sim.destroy()
    sel = amr_tree.leaf_mask(up).repeat(BS ** 2)
    for a, b in ((g1, w1), (g2, w2)):
        err = np.abs(a[sel].astype(np.float64) - b[sel]).max(axis=0) / np.abs(b[sel]).max(axis=0).clip(1e-30)
        assert err.max() <= tol, err
    back = leaf_rows(blocks)
    assert np.abs(b1[back] - u1[back]).max() <= 4 * tol * np.abs(u1[back]).max()


@pytest.mark.parametrize("precision,tol", [(32, 1e-4), (64, 1e-12)])
def test_2d_indicator_matches_oracle(precision, tol):
    """The flag pass on 2D blocks (Loehner's estimator over two axes) against the oracle's
    amr_flag, coarse-fine faces on both levels: same value per leaf, NaN for parents."""
    blocks = quadtree()
    impl = wgsl64 if precision == 64 else wgsl32
    cfg = cfg2(precision)
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, impl.dtype, ndim=2)
    want = impl.amr_flag(orc_prm(cfg), AmrTable(blocks, ROOT, ndim=2), DX0, u1, u2).astype(np.float64)
    sim = AmrSim(cfg, blocks, ROOT, path="passes")
    try:
        sim.set_state(u1, u2)
        ind = sim.indicator()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks)
    assert np.all(np.isnan(ind[~leaf]))
    err = np.abs(ind[leaf] - want[leaf]) / want[leaf]
    print(f"\n[gpu] 2D indicator fp{precision}: range {want[leaf].min():.3g}..{want[leaf].max():.3g}, "
          f"worst rel {err.max():.2e}", flush=True)
    assert want[leaf].max() > 0.05
    assert err.max() <= tol, err


@pytest.mark.parametrize("precision", [32, 64])
def test_2d_bins_match_oracle_cycle(precision):
    """Per-block timesteps on a quadtree: the bins the GPU plans and each cycle's state against
    the oracle's cycle (tests/test_oracle_bins.py pins the cycle)."""
    blocks = quadtree()
    dtype = np.float64 if precision == 64 else np.float32
    w = wgsl64 if precision == 64 else wgsl32
    u1, u2 = amr_tree.blocks_state(blocks, DX0, flow, 1.4, dtype, ndim=2)
    t = AmrTable(blocks, ROOT, ndim=2)
    prm = orc_prm(cfg2(precision))
    sel = leaf_rows(blocks)
    o1, o2 = u1.copy(), u2.copy()
    clk = np.zeros(4, dtype)
    sim = AmrSim(cfg2(precision), blocks, ROOT, path="passes", time_bins=2)
    try:
        sim.set_state(u1, u2)
        for cycle in range(2):
            assert sim.step(1) == 4
            g = sim.read_state()
            bins, work = sim.bins()
            own, _, _ = w.amr_cycle(prm, t, DX0, o1.copy(), o2.copy(), clk.copy(), 2)
            assert np.array_equal(bins, own), (bins.tolist(), own.tolist())
            w.amr_cycle(prm, t, DX0, o1, o2, clk, 2, bins=bins)
            d = l1rel(g, (o1, o2), sel)
            assert max(d.values()) <= (1e-12 if precision == 64 else 1e-5), d
    finally:
        sim.destroy()
    print("2d bins", bins.tolist(), work)
    assert work["updates"] <= work["single"]


def test_2d_refusals():
    from shaderous_amd._lib import ShxError
    blocks = quadtree()
    with pytest.raises(ShxError, match="fused"):
        AmrSim(cfg2(), blocks, ROOT, path="fused")
    with pytest.raises(ShxError, match="coordinate 2 out of range"):
        AmrSim(cfg2(), np.asarray([(0, i, j, 1) for j in range(4) for i in range(4)], np.int32), ROOT, path="passes")
    with pytest.raises(ShxError, match="children"):
        bad = amr_tree.build_tree(ROOT, 1, lambda L, c: c == (0, 0, 0), ndim=2).tolist()
        AmrSim(cfg2(), np.asarray(bad[:-1], np.int32), ROOT, path="passes")
