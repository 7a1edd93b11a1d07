"""Re-meshing of the block-AMR path on an MI355X (shx_amr_indicator, shx_amr_regrid):

- the GPU's new state against the oracle's re-meshing (FP64 to 1e-13, FP32 to 2e-6 per field,
  max-norm relative); the same tree again gives the state bit for bit;
- the indicator (Loehner's estimator, src/amr/flag.wgsl) against the oracle's amr_flag;
- an adaptive run (re-mesh every 4 steps from the indicator): the fused and the pass paths stay
  equal bit for bit, and mass and energy stay conserved to FP32 round-off.
"""
import math

import numpy as np
import pytest

from shaderous_amd import amr_tree
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import SimConfig

oracle = pytest.importorskip("oracle")
from oracle import AmrTable, params, wgsl32, wgsl64  # noqa: E402

pytestmark = pytest.mark.gpu

BS = 16


def blob(x, y, z):
    r2 = x * x + y * y + z * z
    g = np.exp(-r2 / 0.02)
    return {"rho": 1.0 + 0.8 * g, "vx": 0.3 + 0.2 * np.sin(2 * np.pi * y) - 2.0 * x * g,
            "vy": -0.2 * x - 2.0 * y * g, "vz": 0.1 * np.cos(2 * np.pi * z) - 2.0 * z * g,
            "p": 1.0 + 2.0 * np.exp(-r2 / 0.01), "s": (x > 0).astype(float)}


def cfg_for(root, precision=32):
    n = tuple(BS * r for r in root)
    return SimConfig(n=n, ndim=3, dx=1.0 / n[0], gamma=1.4, limiter="mc", isentropic=False, cfl=0.4,
                     precision=precision)


def refined_tree():
    return amr_tree.build_tree((2, 2, 2), 2, lambda L, c: c == (0, 0, 0) or (L == 1 and c == (1, 1, 1)))


def remeshed(blocks):
    ind = np.where(amr_tree.leaf_mask(blocks), 0.0, np.nan)
    for i, b in enumerate(blocks.tolist()):
        if b in ([0, 1, 1, 1], [0, 0, 1, 1]):
            ind[i] = 1.0
    return amr_tree.adapt_tree(blocks, (2, 2, 2), ind, 0.5, 0.1, 2)


def totals(blocks, u1, u2, dx0):
    leaf = amr_tree.leaf_mask(blocks)
    vol = np.repeat((dx0 / (1 << blocks[:, 0].astype(np.int64))) ** 3 * leaf, BS ** 3)
    return np.array([(u1[:, 0].astype(np.float64) * vol).sum(), (u2[:, 0].astype(np.float64) * vol).sum()])


@pytest.mark.parametrize("precision,tol", [(64, 1e-13), (32, 2e-6)])
def test_regrid_matches_oracle(precision, tol):
    print(f"\n[gpu] regrid refined tree fp{precision}", flush=True)
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2), precision)
    impl = wgsl64 if precision == 64 else wgsl32
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma, impl.dtype)
    new = remeshed(blocks)
    k, s, o = amr_tree.regrid_map(blocks, new)
    assert set(np.unique(k)) == {0, 1, 2, 3}
    prm = params(cfg.n, ndim=3, dx=cfg.dx, gamma=cfg.gamma, cfl=cfg.cfl, limiter="mc")
    r1, r2 = impl.amr_regrid(prm, AmrTable(blocks, (2, 2, 2)), u1, u2, k, s, o)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        sim.regrid(new)
        g1, g2 = sim.read_state()
        assert np.array_equal(sim.table(), AmrTable(new, (2, 2, 2)).table())
    finally:
        sim.destroy()
    sel = amr_tree.leaf_mask(new).repeat(BS ** 3)
    for a, b in ((g1, r1), (g2, r2)):
        err = np.abs(a[sel].astype(np.float64) - b[sel]).max(axis=0) / np.abs(b[sel]).max(axis=0).clip(1e-30)
        print(err, flush=True)
        assert err.max() <= tol, err


def test_regrid_to_the_same_tree_is_exact():
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2))
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        sim.step(2)
        a = sim.read_state()
        c = sim.clock()
        sim.regrid(blocks)
        b = sim.read_state()
        assert np.array_equal(sim.clock()[:4], c[:4])
        sim.step(1)
    finally:
        sim.destroy()
    sel = amr_tree.leaf_mask(blocks).repeat(BS ** 3)
    assert np.array_equal(a[0][sel], b[0][sel]) and np.array_equal(a[1][sel], b[1][sel])


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-4)])
def test_indicator_matches_oracle(precision, tol):
    """The flag pass (Loehner's estimator of density and pressure over each leaf's cells, ghosts
    from the assembly) against the oracle's amr_flag on a tree with coarse-fine faces on every
    level: same value per leaf to the precision's rounding, NaN for parents."""
    blocks = refined_tree()
    cfg = cfg_for((2, 2, 2), precision)
    impl = wgsl64 if precision == 64 else wgsl32
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma, impl.dtype)
    prm = params(cfg.n, ndim=3, dx=cfg.dx, gamma=cfg.gamma, cfl=cfg.cfl, limiter="mc")
    want = impl.amr_flag(prm, AmrTable(blocks, (2, 2, 2)), cfg.dx, u1, u2).astype(np.float64)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        ind = sim.indicator()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks)
    assert np.all(np.isnan(ind[~leaf]))
    err = np.abs(ind[leaf] - want[leaf]) / want[leaf]
    print(f"\n[gpu] indicator fp{precision}: leaves {leaf.sum()}, range {want[leaf].min():.3g}..{want[leaf].max():.3g}, "
          f"worst rel {err.max():.2e}", flush=True)
    assert want[leaf].max() > 0.05   # the blob's blocks are flagged well above round-off
    assert err.max() <= tol, err


def test_adaptive_run_paths_agree_and_conserve():
    """The small blast (root 4^3, two levels) re-meshed from the indicator every 4 steps for 12
    steps, in the fused and the pass forms: equal bits, and totals kept to FP32 round-off."""
    blocks, dx0 = amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))
    cfg = cfg_for((4, 4, 4))

    def blast(x, y, z):
        r = np.sqrt((x - 0.03) ** 2 + (y + 0.02) ** 2 + (z - 0.05) ** 2)
        return {"rho": 1.0 + 0.0 * x, "vx": 0.0 * x, "vy": 0.0 * x, "vz": 0.0 * x, "p": np.where(r < 0.2, 100.0, 1.0)}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, blast, cfg.gamma)
    out = {}
    for path in ("fused", "passes"):
        print(f"\n[gpu] adaptive blast {path}", flush=True)
        sim = AmrSim(cfg, blocks, (4, 4, 4), path=path)
        try:
            sim.set_state(u1, u2)
            c0 = totals(sim.blocks, *sim.read_state(), dx0)
            changes = 0
            for _ in range(3):
                sim.step(4)
                changes += sim.adapt(0.05, 0.01, 2)
            s1 = sim.read_state()
            out[path] = (sim.blocks.copy(), s1, changes)
            c1 = totals(sim.blocks, *s1, dx0)
        finally:
            sim.destroy()
        assert np.all(np.abs(c1 - c0) <= 2e-6 * np.abs(c0)), (c1 - c0) / c0
    assert out["fused"][2] > 0
    assert np.array_equal(out["fused"][0], out["passes"][0])
    assert np.array_equal(out["fused"][1][0], out["passes"][1][0]) and np.array_equal(out["fused"][1][1], out["passes"][1][1])


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_auto_regrid_equals_the_callers_cadence(path):
    """The re-mesh cadence inside the steps (layout regrid_every, the reference's REGRID_EVERY):
    a sim stepping 12 steps with regrid_every 4 ends on the same tree, with the same bits, as one
    whose caller steps 4 at a time and adapts from the indicator after each (AmrSim.adapt: the
    Python policy, which tests/test_adapt_native.py holds equal to the library's)."""
    blocks, dx0 = amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))
    cfg = cfg_for((4, 4, 4))

    def blast(x, y, z):
        r = np.sqrt((x - 0.03) ** 2 + (y + 0.02) ** 2 + (z - 0.05) ** 2)
        return {"rho": 1.0 + 0.0 * x, "vx": 0.0 * x, "vy": 0.0 * x, "vz": 0.0 * x, "p": np.where(r < 0.2, 100.0, 1.0)}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, blast, cfg.gamma)
    print(f"\n[gpu] auto regrid {path}", flush=True)
    auto = AmrSim(cfg, blocks, (4, 4, 4), path=path, regrid_every=4, max_level=2, refine_above=0.05,
                  coarsen_below=0.01)
    hand = AmrSim(cfg, blocks, (4, 4, 4), path=path)
    try:
        auto.set_state(u1, u2)
        hand.set_state(u1, u2)
        assert auto.step(10) == 10                       # re-meshes after steps 4 and 8
        assert auto.step(2) == 2                         # and 12
        changes = 0
        for _ in range(3):
            hand.step(4)
            changes += hand.adapt(0.05, 0.01, 2)
        a = auto.read_state()
        h = hand.read_state()
        assert changes > 0 and not np.array_equal(hand.blocks, blocks)
        assert np.array_equal(auto.blocks, hand.blocks) and np.array_equal(auto.table(), hand.table())
        assert np.array_equal(a[0], h[0]) and np.array_equal(a[1], h[1])
        assert np.array_equal(auto.clock(), hand.clock())
        auto.step(3)                                     # within the next period: no re-mesh
        hand.step(3)
        assert np.array_equal(auto.blocks, hand.blocks)
        assert np.array_equal(auto.read_state()[0], hand.read_state()[0])
    finally:
        auto.destroy()
        hand.destroy()
