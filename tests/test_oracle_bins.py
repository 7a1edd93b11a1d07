"""Pins for the oracle's per-block timesteps (src/amr/bins_*.wgsl, amr/mid_on.wgsl, the share
and deposit of amr/update.wgsl:173-290, the cycle of src/amr/amr_sim.ts:1509-1660), on the CPU.

- bins off is the single timestep: a cycle of one substep (kmax 0) is the AMR step, bit for bit;
- the step of a bin is 2^k dt: a cycle in which every block sits in bin 1 is one step at twice
  the CFL number, bit for bit (a dropped or doubled factor 2^k fails it);
- the bin rules, by hand: a periodic chain of blocks one of which is fast ends as the distance
  rule says (0 0 1 2 3 2 1 0), a parent takes its fastest child's bin, a finer block is never
  slower than its coarser neighbour;
- conservation: with several bins live across same-level and coarse-fine faces, the cycle keeps
  mass, momentum and energy to FP64 round-off in a periodic box (a lost deposit, a share of one
  instead of one half, or a slow block's face flux taken whole fails it);
- the cycle follows the single timestep: after the same time, the two differ by the scheme's
  truncation error only.
"""
import numpy as np
import pytest

from oracle import AmrTable, params, wgsl32, wgsl64
from shaderous_amd import amr_tree

BS = 16
GAMMA = 1.4


def bump(x, y, z):
    """A hot spot in the refined corner (the box is [-1/2, 1/2]^3, the refined root block its
    lower octant) over cool gas: the fine leaves there step fastest."""
    r2 = (x + 0.25) ** 2 + (y + 0.25) ** 2 + (z + 0.25) ** 2
    return {"rho": 1.0 + 0.5 * np.exp(-r2 / 0.02), "vx": 0.4 + 0.1 * np.sin(2 * np.pi * y),
            "vy": -0.1 * np.cos(2 * np.pi * x), "vz": 0.05, "p": 0.1 + 4.0 * np.exp(-r2 / 0.005)}


def refined_tree():
    """Root 2^3 blocks, the one at the origin corner refined: 7 coarse leaves, 8 fine ones."""
    return amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))


def setup(dtype, tree=None, root=(2, 2, 2), fn=bump):
    blocks = refined_tree() if tree is None else tree
    dx0 = 1.0 / (BS * root[0])
    u1, u2 = amr_tree.blocks_state(blocks, dx0, fn, GAMMA, dtype)
    t = AmrTable(blocks, root)
    prm = params(tuple(BS * r for r in root), ndim=3, dx=dx0, gamma=GAMMA, cfl=0.4, limiter="mc", isentropic=False)
    return blocks, dx0, u1, u2, t, prm


def totals(blocks, dx0, u1, u2):
    leaf = amr_tree.leaf_mask(blocks)
    vol = np.repeat(np.where(leaf, (dx0 / 2.0 ** blocks[:, 0]) ** 3, 0.0), BS ** 3)
    a = u1.astype(np.float64) * vol[:, None]
    b = u2.astype(np.float64) * vol[:, None]
    return np.array([a[:, 0].sum(), a[:, 1].sum(), a[:, 2].sum(), a[:, 3].sum(), b[:, 0].sum()])


@pytest.mark.parametrize("w", [wgsl32, wgsl64])
def test_one_substep_cycle_is_the_step(w):
    blocks, dx0, u1, u2, t, prm = setup(w.dtype)
    a1, a2 = u1.copy(), u2.copy()
    b1, b2 = u1.copy(), u2.copy()
    ca, cb = np.zeros(4, w.dtype), np.zeros(4, w.dtype)
    for _ in range(2):
        w.amr_step(prm, t, dx0, a1, a2, ca)
        bins, upd, single = w.amr_cycle(prm, t, dx0, b1, b2, cb, 0)
        assert not bins.any() and upd == single == int(amr_tree.leaf_mask(blocks).sum())
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    assert np.array_equal(ca, cb)


@pytest.mark.parametrize("w", [wgsl32, wgsl64])
def test_every_block_in_bin_one_is_a_step_of_twice_the_cfl(w):
    blocks, dx0, u1, u2, t, prm = setup(w.dtype)
    prm2 = params(prm.n, ndim=3, dx=dx0, gamma=GAMMA, cfl=0.8, limiter="mc", isentropic=False)
    a1, a2 = u1.copy(), u2.copy()
    w.amr_step(prm2, t, dx0, a1, a2, np.zeros(4, w.dtype))
    b1, b2 = u1.copy(), u2.copy()
    clk = np.zeros(4, w.dtype)
    bins, upd, single = w.amr_cycle(prm, t, dx0, b1, b2, clk, 1, bins=np.ones(len(blocks), np.int32))
    nleaf = int(amr_tree.leaf_mask(blocks).sum())
    assert upd == nleaf and single == 2 * nleaf
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    assert clk[3] == 1     # the first substep banked at the second


def chain_rates(fast=0, n=8):
    r = np.full(n, 1.0 / 8.0)
    r[fast] = 1.0
    return r


def test_bin_rules_on_a_chain():
    """Eight root blocks along x, periodic: block 0 eight times faster than the rest. Its own
    bins are 0 and 3; the dilation brings its two neighbours to 0; one bin per block beyond."""
    blocks = np.array([(0, i, 0, 0) for i in range(8)], np.int32)
    t = AmrTable(blocks, (8, 1, 1))
    prm = params((128, 16, 16), ndim=3, dx=1.0 / 128, gamma=GAMMA, cfl=0.4)
    for w in (wgsl32, wgsl64):
        dt = w.dtype.type(0.4 / 1.0)          # cfl / the fastest block's rate
        k = w.amr_bins(prm, t, 3, chain_rates(), dt)
        assert k.tolist() == [0, 0, 1, 2, 3, 2, 1, 0]
        # with at most 2 bins the far blocks stop at 2
        assert w.amr_bins(prm, t, 2, chain_rates(), dt).tolist() == [0, 0, 1, 2, 2, 2, 1, 0]
        # a held clock (dt 0): every block in the slowest bin
        assert w.amr_bins(prm, t, 3, chain_rates(), w.dtype.type(0)).tolist() == [3] * 8
        # floor(log2): a block a little slower than twice the fastest's step still bins 0
        r = np.full(8, 1.0 / 1.9)
        r[0] = 1.0
        assert w.amr_bins(prm, t, 3, r, dt).tolist() == [0] * 8


def test_bin_rules_across_levels():
    """A refined root block whose fine leaves are all slow, beside fast coarse ones: a fine leaf
    is never slower than its coarser neighbour; the parent takes its fastest child's bin."""
    blocks = refined_tree()
    t = AmrTable(blocks, (2, 2, 2))
    tab = t.table()
    leaf = amr_tree.leaf_mask(blocks)
    fine = blocks[:, 0] == 1
    prm = params((32, 32, 32), ndim=3, dx=1.0 / 32, gamma=GAMMA, cfl=0.4)
    r = np.where(fine, 1.0 / 16.0, 1.0)            # coarse leaves fastest (bin 0), fine ones bin 4 alone
    r[~leaf] = 0.0
    k = wgsl64.amr_bins(prm, t, 3, r, np.float64(0.4))
    parent = np.flatnonzero(~leaf)[0]
    for s in np.flatnonzero(leaf & fine):
        coarse = [tab[s, 20 + f] for f in range(6) if tab[s, 6 + f] == -1]
        assert coarse, "every fine leaf of the refined corner touches a coarse one (periodic box)"
        assert k[s] <= min(k[c] for c in coarse) == 0
    assert k[parent] == min(k[c] for c in tab[parent, 12:20] if c >= 0)
    # the other way: fast fine leaves, slow coarse ones. The parent stands for its children (bin
    # 0); a coarse leaf is one bin per face hop from it: on the periodic 2^3 root grid, the number
    # of its block coordinates that are 1 (the refined block is at the origin)
    r = np.where(fine, 1.0, 1.0 / 16.0)
    r[~leaf] = 0.0
    k = wgsl64.amr_bins(prm, t, 3, r, np.float64(0.4))
    assert (k[leaf & fine] == 0).all()
    assert k[parent] == 0
    for s in np.flatnonzero(leaf & ~fine):
        assert k[s] == min(3, int(blocks[s, 1:].sum())), (blocks[s], k[s])


def live_bins(blocks):
    """Bins that follow the rules with jumps live on every kind of face: fine leaves and their
    parent 0, each coarse leaf one bin per face hop from the refined block (1 or 2: at most 2
    with kmax 2), so same-level faces between bins 1 and 2 and coarse-fine faces between 0 and 1."""
    leaf = amr_tree.leaf_mask(blocks)
    hops = np.minimum(blocks[:, 1:].sum(1), 2)
    return np.where((blocks[:, 0] == 1) | ~leaf, 0, hops).astype(np.int32)


def test_cycle_conserves_with_live_bins():
    blocks, dx0, u1, u2, t, prm = setup(np.float64)
    t0 = totals(blocks, dx0, u1, u2)
    clk = np.zeros(4)
    bins, upd, single = wgsl64.amr_cycle(prm, t, dx0, u1, u2, clk, 2, bins=live_bins(blocks))
    assert upd < single
    t1 = totals(blocks, dx0, u1, u2)
    assert np.abs(t1 - t0).max() < 1e-13 * np.abs(t0).max(), t1 - t0
    # two bins of the cycle's own, over several cycles: mass and energy to round-off still
    for _ in range(3):
        bins, upd, single = wgsl64.amr_cycle(prm, t, dx0, u1, u2, clk, 2)
    t2 = totals(blocks, dx0, u1, u2)
    assert np.abs(t2 - t0).max() < 1e-13 * np.abs(t0).max(), t2 - t0


def test_cycle_uses_several_bins_and_saves_updates():
    """The cycle's own bins on the bump: the fine leaves over the hot spot are fastest, the far
    coarse leaves slower; fewer block updates than single steps."""
    blocks, dx0, u1, u2, t, prm = setup(np.float64)
    bins, upd, single = wgsl64.amr_cycle(prm, t, dx0, u1, u2, np.zeros(4), 2)
    leaf = amr_tree.leaf_mask(blocks)
    assert len(set(bins[leaf].tolist())) >= 2
    assert upd < single


def test_cycle_follows_the_single_timestep():
    """Smooth advection, the coarse leaves one bin slower than the fine ones: one cycle (fine
    leaves two steps of dt, coarse ones one of 2 dt, a deposit at the coarse side's midpoint)
    against two single steps over the same time. The fine leaves' ghosts from a coarse leaf half
    way through its step come from its half-step prediction (amr/mid_on.wgsl): the two runs then
    differ by ~3e-7 in L1 of density; ghosts taken from the coarse leaf's end-of-step state
    instead would leave 8e-4."""
    blocks = refined_tree()
    dx0 = 1.0 / 32

    def wave(x, y, z):
        return {"rho": 2.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), "vx": 1.0, "vy": 0.3, "vz": 0.2,
                "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, wave, GAMMA, np.float64)
    t = AmrTable(blocks, (2, 2, 2))
    prm = params((32, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, cfl=0.4, limiter="mc", isentropic=False)
    leaf = amr_tree.leaf_mask(blocks)
    bins = np.where((blocks[:, 0] == 1) | ~leaf, 0, 1).astype(np.int32)
    a1, a2 = u1.copy(), u2.copy()
    _, upd, single = wgsl64.amr_cycle(prm, t, dx0, a1, a2, np.zeros(4), 1, bins=bins)
    # fine 2 x 8, coarse 7, and deposits for the 3 coarse leaves sharing a face with the fine ones
    assert upd == 2 * 8 + 7 + 3 and single == 30
    b1, b2 = u1.copy(), u2.copy()
    c = np.zeros(4)
    for _ in range(2):
        wgsl64.amr_cycle(prm, t, dx0, b1, b2, c, 0, bins=np.zeros(len(blocks), np.int32))
    fine = ((blocks[:, 0] == 1) & leaf).repeat(BS ** 3)
    coarse = ((blocks[:, 0] == 0) & leaf).repeat(BS ** 3)
    assert np.abs(a1[fine, 0] - b1[fine, 0]).mean() < 1e-5
    assert np.abs(a1[coarse, 0] - b1[coarse, 0]).mean() < 1e-4


def test_same_level_bins_follow_single_steps():
    """Smooth advection on 4 x 2 x 2 root blocks, the half at x > 0 one bin slower: one cycle
    against two single steps. The fast blocks' ghosts from a slow neighbour half way through its
    step are its half-step prediction, predicted again in place as ghosts beside a coarser block
    are (amr/predict.wgsl:193-197): the fast half differs by ~2e-7 in L1 of density. Ghosts from
    the slow block's end-of-step state leave 9e-4, and predicting them at their source from the
    slow block's start-of-step primitives 8e-6."""
    root = (4, 2, 2)
    blocks = np.array([(0, x, y, z) for z in range(2) for y in range(2) for x in range(4)], np.int32)
    dx0 = 1.0 / 64

    def wave(x, y, z):
        return {"rho": 2.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), "vx": 1.0, "vy": 0.3, "vz": 0.2,
                "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, wave, GAMMA, np.float64)
    t = AmrTable(blocks, root)
    prm = params((64, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, cfl=0.4, limiter="mc", isentropic=False)
    bins = (blocks[:, 1] >= 2).astype(np.int32)
    a1, a2 = u1.copy(), u2.copy()
    _, upd, single = wgsl64.amr_cycle(prm, t, dx0, a1, a2, np.zeros(4), 1, bins=bins)
    # every slow block has a fast neighbour (across x = 0 or the periodic wrap): all get deposits
    assert upd == 16 + 8 + 8 and single == 32
    b1, b2 = u1.copy(), u2.copy()
    for _ in range(2):
        wgsl64.amr_cycle(prm, t, dx0, b1, b2, np.zeros(4), 0, bins=np.zeros(len(blocks), np.int32))
    fast, slow = (bins == 0).repeat(BS ** 3), (bins == 1).repeat(BS ** 3)
    assert np.abs(a1[fast, 0] - b1[fast, 0]).mean() < 1e-6
    assert np.abs(a1[slow, 0] - b1[slow, 0]).mean() < 1e-4
