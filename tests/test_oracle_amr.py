"""Pins for the block-AMR oracle (the src/amr WGSL passes and the tree table), on the CPU.

- tree invariants the reference's tests check (tests/unit/tree.test.ts:26-391): 2:1 balance,
  reciprocal neighbours, every point of the domain covered by exactly one leaf;
- `amr-unrefined` (src/validation/tests.ts:1202-1229): a block grid with no refinement steps like
  the uniform grid (the uniform WGSL oracle, itself pinned by tests/test_oracle.py) within 1e-5;
- the prolongation is conservative (the mean of the 8 children is the parent), reproduces linear
  fields and keeps rho, E, e at least half the coarse cell's;
- conservation across refinement boundaries to round-off (the coarse-fine flux correction), in
  FP64, and FP32 against FP64.
"""
import numpy as np
import pytest

import oracle
from oracle import AmrTable, params, wgsl32, wgsl64
from shaderous_amd import amr_tree, problems

BS = 16


def blocks_from_uniform(u, root):
    """(nz, ny, nx, 4) -> (nblocks * 4096, 4) in the root-grid block order (z, y, x)."""
    out = []
    for bz in range(root[2]):
        for by in range(root[1]):
            for bx in range(root[0]):
                out.append(u[BS * bz:BS * bz + BS, BS * by:BS * by + BS, BS * bx:BS * bx + BS].reshape(-1, 4))
    return np.ascontiguousarray(np.concatenate(out))


def root_blocks(root):
    return np.asarray([(0, x, y, z) for z in range(root[2]) for y in range(root[1]) for x in range(root[0])], np.int32)


def blob(x, y, z):
    r2 = x * x + y * y + z * z
    return {"rho": 1.0 + 0.8 * np.exp(-r2 / 0.02), "vx": 0.3 + 0.2 * np.sin(2 * np.pi * y),
            "vy": -0.2 * x, "vz": 0.1 * np.cos(2 * np.pi * z), "p": 1.0 + 2.0 * np.exp(-r2 / 0.01)}


# ---------------- tree ----------------

def small_blast_tree():
    return amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))


def test_tree_invariants():
    blocks, dx0 = small_blast_tree()
    t = AmrTable(blocks, (4, 4, 4))
    tab = t.table()
    leaf = amr_tree.leaf_mask(blocks)
    assert np.array_equal((tab[:, 4] & 2) != 0, leaf)
    # leaf volumes tile the unit box
    assert abs(sum(8.0 ** -int(b[0]) for b in blocks[leaf]) / 64 - 1.0) < 1e-12
    # points fall in exactly one leaf
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (500, 3))
    for p in pts:
        hits = 0
        for b in blocks[leaf]:
            s = 0.25 / (1 << int(b[0]))
            lo = b[1:] * s
            hits += bool(np.all(p >= lo) and np.all(p < lo + s))
        assert hits == 1
    for s, r in enumerate(tab):
        assert r[5] == -1 if r[0] == 0 else tab[r[5], 0] == r[0] - 1
        for o in range(8):
            if r[12 + o] >= 0:
                assert tab[r[12 + o], 5] == s
        for f in range(6):
            n = r[6 + f]
            if n >= 0:
                assert tab[n, 6 + (f ^ 1)] == s          # reciprocal
                assert tab[n, 0] == r[0]
            else:
                assert n == -1                            # periodic box: no boundary codes
                c = r[20 + f]
                assert tab[c, 0] == r[0] - 1 and (tab[c, 4] & 2)
                assert r[0] >= 1
    # 2:1 balance across faces: a leaf's face neighbour region holds leaves of level L-1..L+1
    for s, r in enumerate(tab):
        if not leaf[s]:
            continue
        for f in range(6):
            n = r[6 + f]
            if n >= 0 and not leaf[n]:
                for o in range(8):
                    ch = tab[n, 12 + o]
                    assert ch >= 0


def test_tree_rejects_unbalanced():
    blocks = [(0, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    blocks += [(1, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    blocks += [(2, x, y, z) for z in range(2) for y in range(2) for x in range(2)]
    with pytest.raises(ValueError):
        AmrTable(np.asarray(blocks), (2, 2, 2))
    # the builder refines the neighbours instead
    b = amr_tree.build_tree((2, 2, 2), 2, lambda L, c: c == (0, 0, 0))
    AmrTable(b, (2, 2, 2))


# ---------------- prolongation (src/amr/cons.wgsl:57-124) ----------------

def random_stencil(rng):
    v = np.empty((7, 8))
    for k in range(7):
        rho = rng.uniform(0.5, 2.0)
        vel = rng.uniform(-1, 1, 3)
        p = rng.uniform(0.5, 2.0)
        e = p / 0.4
        v[k] = [rho, *(rho * vel), e + 0.5 * rho * vel @ vel, e, rho * np.log(p / rho ** 1.4), 0.0]
    return v


def test_prolong_conservative():
    rng = np.random.default_rng(7)
    prm = params((16, 16, 16), limiter="mc")
    for trial in range(40):
        st = random_stencil(rng)
        ok = [1] + list(rng.integers(0, 2, 6)) if trial % 3 == 0 else [1] * 7
        kids = np.asarray([wgsl64.amr_prolong(prm, st, ok, (i, j, k)) for k in (0, 1) for j in (0, 1) for i in (0, 1)])
        assert np.allclose(kids.mean(axis=0)[:7], st[0, :7], rtol=0, atol=1e-14 * np.abs(st[0, :7]).max())
        # rho, E, e never below half the coarse cell's (the joint scale-back)
        for q in (0, 4, 5):
            assert kids[:, q].min() >= 0.5 * st[0, q] - 1e-14


def test_prolong_reproduces_linear():
    prm = params((16, 16, 16), limiter="mc")
    g = np.array([0.1, -0.05, 0.02])         # density gradient per coarse cell
    v = np.array([0.2, 0.1, -0.3])           # uniform velocity
    p = 1.0
    st = np.empty((7, 8))
    offs = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    for k, o in enumerate(offs):
        rho = 1.0 + g @ np.asarray(o, float)
        st[k] = [rho, *(rho * v), p / 0.4 + 0.5 * rho * v @ v, p / 0.4, 0.0, 0.0]
    for par in ((0, 0, 0), (1, 0, 1), (1, 1, 1)):
        out = wgsl64.amr_prolong(prm, st, [1] * 7, par)
        x = (np.asarray(par) - 0.5) * 0.5
        assert abs(out[0] - (1.0 + g @ x)) < 1e-14
        assert np.allclose(out[1:4], (1.0 + g @ x) * v, atol=1e-14)     # momentum slopes by the product rule
        assert abs(out[5] - p / 0.4) < 1e-14                            # flat e stays flat
    # a one-sided stencil drops the axis's slope
    out = wgsl64.amr_prolong(prm, st, [1, 0, 1, 1, 1, 1, 1], (1, 0, 0))
    assert abs(out[0] - (1.0 + g[1] * -0.25 + g[2] * -0.25)) < 1e-14


# ---------------- unrefined blocks == uniform grid (amr-unrefined) ----------------

@pytest.mark.parametrize("impl,tol", [(wgsl64, 1e-12), (wgsl32, 1e-5)])
@pytest.mark.parametrize("bcs", ["periodic", "walls"])
def test_amr_unrefined_matches_uniform(impl, tol, bcs):
    root = (2, 2, 2)
    n = (32, 32, 32)
    bc_lo, bc_hi = ((0, 0, 0), (0, 0, 0)) if bcs == "periodic" else ((1, 2, 0), (2, 1, 0))
    prm = params(n, ndim=3, gamma=1.4, cfl=0.4, limiter="mc", bc_lo=bc_lo, bc_hi=bc_hi)
    u1, u2 = problems.cons_from_prims(blob(*problems.centers(n, 1.0)), 1.4, impl.dtype)
    b1 = blocks_from_uniform(u1.reshape(32, 32, 32, 4), root)
    b2 = blocks_from_uniform(u2.reshape(32, 32, 32, 4), root)
    t = AmrTable(root_blocks(root), root, bc_lo, bc_hi)
    cu = np.zeros(4, impl.dtype)
    ca = np.zeros(4, impl.dtype)
    for _ in range(3):
        impl.step(prm, u1, u2, cu)
        impl.amr_step(prm, t, 1.0 / 32, b1, b2, ca)
        assert abs(ca[0] - cu[0]) <= tol * cu[0]
        ref = blocks_from_uniform(u1.reshape(32, 32, 32, 4), root)
        rel = np.abs(b1 - ref).max(axis=0) / np.abs(ref).max(axis=0).clip(1e-30)
        assert rel.max() <= tol, rel


# ---------------- refined trees: conservation and precision ----------------

def refined_tree():
    # root 2^3; one root block refined twice at a corner: levels 0..2, balanced by the builder
    return amr_tree.build_tree((2, 2, 2), 2, lambda L, c: c == (0, 0, 0) or (L == 1 and c == (1, 1, 1)))


def amr_state(blocks, dtype):
    return amr_tree.blocks_state(blocks, 1.0 / 32, blob, 1.4, dtype)


def totals(blocks, u1, u2):
    leaf = amr_tree.leaf_mask(blocks)
    vol = np.repeat((1.0 / 32 / (1 << blocks[:, 0].astype(np.int64))) ** 3 * leaf, BS ** 3)
    return np.concatenate([(u1.astype(np.float64) * vol[:, None]).sum(0), [(u2[:, 0].astype(np.float64) * vol).sum()]])


@pytest.mark.parametrize("compensated", [True, False])
def test_amr_conservation_fp64(compensated):
    blocks = refined_tree()
    assert len(np.unique(blocks[:, 0])) == 3
    t = AmrTable(blocks, (2, 2, 2))
    u1, u2 = amr_state(blocks, np.float64)
    prm = params((32, 32, 32), ndim=3, gamma=1.4, cfl=0.4, limiter="mc")
    c0 = totals(blocks, u1, u2)
    clk = np.zeros(4)
    for _ in range(3):
        wgsl64.amr_step(prm, t, 1.0 / 32, u1, u2, clk, compensated)
    c1 = totals(blocks, u1, u2)
    scale = np.abs(c0) + np.array([0, 1, 1, 1, 0]) * c0[0]
    assert np.all(np.abs(c1 - c0) <= 1e-12 * scale), (c1 - c0) / scale   # ~1e-13: sums of 50k cells
    assert np.all(u1[:, 0][amr_tree.leaf_mask(blocks).repeat(BS ** 3)] > 0)


def test_amr_fp32_tracks_fp64():
    blocks = refined_tree()
    t = AmrTable(blocks, (2, 2, 2))
    prm = params((32, 32, 32), ndim=3, gamma=1.4, cfl=0.4, limiter="mc")
    a1, a2 = amr_state(blocks, np.float32)
    d1, d2 = amr_state(blocks, np.float64)
    ca, cd = np.zeros(4, np.float32), np.zeros(4)
    for _ in range(3):
        wgsl32.amr_step(prm, t, 1.0 / 32, a1, a2, ca)
        wgsl64.amr_step(prm, t, 1.0 / 32, d1, d2, cd)
    leaf = amr_tree.leaf_mask(blocks).repeat(BS ** 3)
    for q in range(4):
        l1 = np.abs(a1[leaf, q] - d1[leaf, q]).sum() / np.abs(d1[leaf, q]).sum()
        assert l1 < 1e-5, (q, l1)
    # FP32 conserves to its own round-off
    c0 = totals(blocks, *amr_state(blocks, np.float32))
    c1 = totals(blocks, a1, a2)
    assert abs(c1[0] - c0[0]) / c0[0] < 1e-6
    assert abs(c1[4] - c0[4]) / c0[4] < 1e-6


def test_amr_ghosts_of_linear_field():
    """Ghost texels filled across refinement boundaries hold a linear density field exactly: from
    a coarser neighbour by prolongation (MC slopes of a linear field are its gradient), from a
    finer one by restriction (the mean of 8 children is the value at the centre)."""
    root = (4, 4, 4)
    blocks = amr_tree.build_tree(root, 1, lambda L, c: c == (1, 1, 1) or c == (2, 1, 1))
    t = AmrTable(blocks, root)
    tab = t.table()
    dx0 = 1.0 / 64
    g = np.array([0.3, -0.2, 0.1])
    lin = lambda x, y, z: {"rho": 2.0 + g[0] * x + g[1] * y + g[2] * z, "vx": 0.1, "vy": -0.2, "vz": 0.05, "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, lin, 1.4, np.float64)
    prm = params((64, 64, 64), ndim=3, gamma=1.4, limiter="mc")
    W, _ = wgsl64.amr_assemble(prm, t, dx0, u1, u2)
    checked = {"coarse": 0, "fine": 0}
    for s, r in enumerate(tab):
        if not (r[4] & 2):
            continue
        L = r[0]
        dx = dx0 / (1 << L)
        for f in range(6):
            n = r[6 + f]
            kind = "coarse" if n == -1 else ("fine" if n >= 0 and not (tab[n, 4] & 2) else None)
            if kind is None:
                continue
            # the block must sit away from the periodic wrap, where the field is not linear
            lo = (r[1:4] * 16 - 2) * dx - 0.5
            hi = (r[1:4] * 16 + 18) * dx - 0.5
            if np.any(lo < -0.5 + 2 * dx0) or np.any(hi > 0.5 - 2 * dx0):
                continue
            axis, up = f >> 1, f & 1
            for layer in (0, 1):
                idx = [slice(2, 18)] * 3
                idx[2 - axis] = 18 + layer if up else 1 - layer
                x = (np.arange(-2, 18) + 0.5 + 16 * r[1]) * dx - 0.5
                y = (np.arange(-2, 18) + 0.5 + 16 * r[2]) * dx - 0.5
                z = (np.arange(-2, 18) + 0.5 + 16 * r[3]) * dx - 0.5
                zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
                want = 2.0 + g[0] * xx + g[1] * yy + g[2] * zz
                got = W[s][..., 0][tuple(idx)]
                assert np.allclose(got, want[tuple(idx)], rtol=0, atol=1e-13), (s, f, layer, kind)
                assert np.allclose(W[s][..., 1][tuple(idx)], 0.1, atol=1e-13)
                checked[kind] += 1
    assert checked["coarse"] > 0 and checked["fine"] > 0, checked


# ---------------- re-meshing ----------------

def remeshed(blocks):
    """The refined tree re-meshed once: the level-1 parent at the corner merged back into a
    leaf, two level-0 leaves split (within one level of the old tree everywhere)."""
    ind = np.where(amr_tree.leaf_mask(blocks), 0.0, np.nan)
    for i, b in enumerate(blocks.tolist()):
        if b in ([0, 1, 1, 1], [0, 0, 1, 1]):
            ind[i] = 1.0
        if b[0] == 2:
            ind[i] = 0.0
    return amr_tree.adapt_tree(blocks, (2, 2, 2), ind, 0.5, 0.1, 2)


def test_regrid_map_kinds():
    blocks = refined_tree()
    k, s, o = amr_tree.regrid_map(blocks, blocks)
    leaf = amr_tree.leaf_mask(blocks)
    assert np.all(k[leaf] == 1) and np.all(k[~leaf] == 0) and np.array_equal(s[leaf], np.flatnonzero(leaf))
    new = remeshed(blocks)
    k, s, o = amr_tree.regrid_map(blocks, new)
    assert set(np.unique(k)) == {0, 1, 2, 3}
    for j in np.flatnonzero(k == 2):                     # a prolonged child: its parent is an old leaf
        L, x, y, z = new[j]
        assert tuple(blocks[s[j]]) == (L - 1, x >> 1, y >> 1, z >> 1) and o[j] == (x & 1) | (y & 1) << 1 | (z & 1) << 2
    for j in np.flatnonzero(k == 3):                     # a merged block was an old parent
        assert tuple(blocks[s[j]]) == tuple(new[j]) and not leaf[s[j]]
    # two levels at once is refused
    with pytest.raises(ValueError):
        amr_tree.regrid_map(root_blocks((2, 2, 2)), blocks)


def test_adapt_tree_stays_within_a_level():
    blocks, _ = small_blast_tree()
    rng = np.random.default_rng(3)
    leaf = amr_tree.leaf_mask(blocks)
    for _ in range(3):
        ind = np.where(leaf, rng.uniform(0, 1, len(blocks)), np.nan)
        new = amr_tree.adapt_tree(blocks, (4, 4, 4), ind, 0.8, 0.3, 3)
        AmrTable(new, (4, 4, 4))                          # 2:1 balanced (raises otherwise)
        amr_tree.regrid_map(blocks, new)                  # within one level (raises otherwise)


@pytest.mark.parametrize("impl,tol", [(wgsl64, 1e-13), (wgsl32, 2e-6)])
def test_regrid_conserves_and_round_trips(impl, tol):
    """Prolongation and restriction keep the totals to round-off; prolonging every leaf and
    restricting back gives the state again (the children's mean is the coarse cell, energy
    included)."""
    blocks = refined_tree()
    t = AmrTable(blocks, (2, 2, 2))
    u1, u2 = amr_state(blocks, impl.dtype)
    prm = params((32, 32, 32), ndim=3, gamma=1.4, cfl=0.4, limiter="mc")
    new = remeshed(blocks)
    k, s, o = amr_tree.regrid_map(blocks, new)
    n1, n2 = impl.amr_regrid(prm, t, u1, u2, k, s, o)
    c0, c1 = totals(blocks, u1, u2), totals(new, n1, n2)
    scale = np.abs(c0) + np.array([0, 1, 1, 1, 0]) * c0[0]
    assert np.all(np.abs(c1 - c0) <= tol * scale), (c1 - c0) / scale
    # split every leaf below level 2, then merge the new parents back
    leaf = amr_tree.leaf_mask(blocks)
    up = amr_tree.adapt_tree(blocks, (2, 2, 2), np.where(leaf, 1.0, np.nan), 0.5, -1.0, 2)
    k, s, o = amr_tree.regrid_map(blocks, up)
    assert (k == 2).sum() > 0
    f1, f2 = impl.amr_regrid(prm, t, u1, u2, k, s, o)
    tu = AmrTable(up, (2, 2, 2))
    kb, sb, ob = amr_tree.regrid_map(up, blocks)
    b1, b2 = impl.amr_regrid(prm, tu, f1, f2, kb, sb, ob)
    sel = leaf.repeat(BS ** 3)
    for a, b in ((b1, u1), (b2, u2)):
        err = np.abs(a[sel].astype(np.float64) - b[sel]).max(axis=0) / np.abs(b[sel]).max(axis=0).clip(1e-30)
        assert err.max() <= 4 * tol, err


def test_regrid_prolongs_linear_fields_exactly():
    """A linear density and momentum field on a uniform root grid: every child of a split
    block takes the linear field's value at its own centre (slopes within the limiter's
    range), and restricting back returns it."""
    root = (2, 2, 2)
    blocks = root_blocks(root)
    lin = lambda x, y, z: {"rho": 2.0 + 0.5 * x + 0.25 * y - 0.125 * z, "vx": 0.0 * x, "vy": 0.0 * x, "vz": 0.0 * x,
                           "p": 1.0 + 0.0 * x}
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, lin, 1.4, np.float64)
    t = AmrTable(blocks, root)
    prm = params((32, 32, 32), ndim=3, gamma=1.4, cfl=0.4, limiter="mc")
    up = amr_tree.adapt_tree(blocks, root, np.ones(len(blocks)), 0.5, -1.0, 1)
    k, s, o = amr_tree.regrid_map(blocks, up)
    n1, _ = wgsl64.amr_regrid(prm, t, u1, u2, k, s, o)
    e1, _ = amr_tree.blocks_state(up, 1.0 / 32, lin, 1.4, np.float64)
    # away from the periodic wrap (where the linear field jumps) the children are exact
    inner = np.zeros(len(up) * BS ** 3, bool)
    for j, (L, x, y, z) in enumerate(up.tolist()):
        if L == 1:
            kk, jj, ii = np.meshgrid(np.arange(BS), np.arange(BS), np.arange(BS), indexing="ij")
            g = [BS * c + a for c, a in ((x, ii), (y, jj), (z, kk))]
            ok = np.all([(gi >= 2) & (gi < 64 - 2) for gi in g], axis=0)
            inner[j * BS ** 3:(j + 1) * BS ** 3] = ok.ravel()
    assert inner.sum() > 0
    assert np.abs(n1[inner, 0] - e1[inner, 0]).max() < 1e-13
