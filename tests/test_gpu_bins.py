"""Per-block timesteps on the GPU (shx_amr layout time_bins; src/amr/bins_*.wgsl, the cycle of
src/amr/amr_sim.ts:1509-1660) through the C ABI, against the oracle's cycle (tests/
test_oracle_bins.py pins it): the bins the GPU plans equal the oracle's, each cycle's state
matches the oracle's cycle run with those bins (FP32 to the L1 bar 1e-5 per field, FP64 to
1e-12), the block updates are counted alike, totals are conserved to round-off (FP64), and the
AMR blast does the fewer block updates its bins allow (1.35x)."""
import numpy as np
import pytest

from oracle import AmrTable, params, wgsl32, wgsl64
from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import SimConfig
from test_oracle_bins import BS, GAMMA, refined_tree, totals

pytestmark = pytest.mark.gpu


def flow(x, y, z):
    """The hot spot of the oracle's pins in a compressive flow. The shock detector compares the
    velocity convergence of neighbouring cells with zero, and at the rim of a face beside a
    coarser block one of them is an empty edge texel: a divergence-free or uniform velocity
    leaves that convergence at the rounding level, whose sign then differs between the GPU's
    v = m * (1 / rho) and the oracle's m / rho (each flip switches a face between HLLC and HLL).
    Velocities bounded away from zero with a divergence keep every such sign clear."""
    r2 = (x + 0.25) ** 2 + (y + 0.25) ** 2 + (z + 0.25) ** 2
    return {"rho": 1.0 + 0.5 * np.exp(-r2 / 0.02), "vx": 0.4 + 0.1 * np.sin(2 * np.pi * x),
            "vy": 0.3 + 0.1 * np.sin(2 * np.pi * y + 1.0), "vz": 0.25 + 0.05 * np.sin(2 * np.pi * z + 2.0),
            "p": 0.1 + 4.0 * np.exp(-r2 / 0.005)}


def l1(a, b, sel):
    out = []
    for x, y in zip(a, b):
        for c in range(4):
            den = np.abs(y[sel, c]).sum()
            if den > 0:
                out.append(np.abs(x[sel, c].astype(np.float64) - y[sel, c]).sum() / den)
    return max(out)


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("path", ["passes", "levels"])
def test_cycles_match_the_oracle(precision, path):
    dtype = np.float64 if precision == 64 else np.float32
    w = wgsl64 if precision == 64 else wgsl32
    blocks = refined_tree()
    dx0 = 1.0 / 32
    u1, u2 = amr_tree.blocks_state(blocks, dx0, flow, GAMMA, dtype)
    t = AmrTable(blocks, (2, 2, 2))
    prm = params((32, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, cfl=0.4, limiter="mc", isentropic=False)
    cfg = SimConfig(n=(32, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, limiter="mc", isentropic=False, cfl=0.4,
                    precision=precision)
    leaf = amr_tree.leaf_mask(blocks)
    sel = leaf.repeat(BS ** 3)
    o1, o2 = u1.copy(), u2.copy()
    clk = np.zeros(4, dtype)
    sim = AmrSim(cfg, blocks, (2, 2, 2), path=path, time_bins=2)
    try:
        sim.set_state(u1, u2)
        upd = single = 0
        for cycle in range(2):
            taken = sim.step(1)
            assert taken == 4                                  # whole cycles of 2^2 substeps
            g1, g2 = sim.read_state()
            bins, work = sim.bins()
            own, _, _ = w.amr_cycle(prm, t, dx0, o1.copy(), o2.copy(), clk.copy(), 2)
            print(precision, path, cycle, "bins", bins.tolist(), "oracle", own.tolist())
            assert np.array_equal(bins, own)
            _, u, sgl = w.amr_cycle(prm, t, dx0, o1, o2, clk, 2, bins=bins)
            upd += u
            single += sgl
            assert work == {"cycles": cycle + 1, "updates": upd, "single": single}
            err = l1((g1, g2), (o1, o2), sel)
            print("  L1", err)
            assert err < (1e-12 if precision == 64 else 1e-5)
            ck = sim.clock()
            assert ck[3] == clk[3] and np.isclose(ck[1], clk[1], rtol=1e-6, atol=0)
        assert len(set(bins[leaf].tolist())) >= 2 and upd < single
    finally:
        sim.destroy()


def test_cycles_conserve_in_fp64():
    blocks = refined_tree()
    dx0 = 1.0 / 32
    u1, u2 = amr_tree.blocks_state(blocks, dx0, flow, GAMMA, np.float64)
    cfg = SimConfig(n=(32, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, limiter="mc", isentropic=False, cfl=0.4, precision=64)
    t0 = totals(blocks, dx0, u1, u2)
    sim = AmrSim(cfg, blocks, (2, 2, 2), path="passes", time_bins=3)
    try:
        sim.set_state(u1, u2)
        sim.step(24)
        g1, g2 = sim.read_state()
        bins, work = sim.bins()
    finally:
        sim.destroy()
    t1 = totals(blocks, dx0, g1, g2)
    assert work["cycles"] == 3 and work["updates"] < work["single"]
    assert np.abs(t1 - t0).max() < 1e-12 * np.abs(t0).max(), t1 - t0


def test_blast_saves_block_updates():
    """The AMR blast (5024 blocks, 3 levels) with cycles of 8 substeps saves what the bin rules
    allow there: 1.35x fewer block updates than single steps (44% of the leaves, the hot sphere's
    finest blocks and their face neighbours, step every substep; the rules bring any block at most
    one bin above a face neighbour and never above a coarser one)."""
    blocks, dx0, prims, centre = problems.amr_blast(seed=7)
    u1, u2 = amr_tree.blocks_state(blocks, dx0, prims, 1.4)
    cfg = SimConfig(n=(128, 128, 128), ndim=3, dx=dx0, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, (8, 8, 8), path="passes", time_bins=3)
    try:
        sim.set_state(u1, u2)
        d0 = sim.diagnostics()
        sim.step(16)
        bins, work = sim.bins()
        d1 = sim.diagnostics()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks)
    hist = np.bincount(bins[leaf], minlength=4)
    print("bins", hist.tolist(), work, "saving", work["single"] / work["updates"])
    assert work["cycles"] == 2
    assert work["single"] >= 1.3 * work["updates"]
    assert abs(d1.mass - d0.mass) < 1e-6 * d0.mass
