"""Pins for the 2D block-AMR oracle (quadtrees of 16^2-cell blocks: src/amr/tree.ts:138-175 with
ndim 2, the is2d() branches of src/amr/{grid,cons,assemble,predict,update,regrid}.wgsl), on the CPU.

- the quadtree table: four children, z faces naming the block itself (tree.ts writeRow
  :662-666), reciprocal neighbours, leaves tiling the square;
- an unrefined quadtree steps like the uniform 2D grid (the uniform oracle, pinned by
  tests/test_oracle.py);
- a refined quadtree steps like its z extrusion, the octree whose every block is a column of
  like blocks along z, run by the 3D restatement (pinned by tests/test_oracle_amr.py): the 2D
  branches (four children per mean, two fine faces per coarse face, no z neighbours in the
  prolongation stencil, the cell for its own z neighbours in the predictor) must give what the 3D
  passes give on a state that does not vary along z, where the z slopes are zero and the z faces
  carry equal fluxes on both sides;
- conservation to round-off in FP64, ghosts of a linear field exact across refinement
  boundaries, and the re-meshing's conservation and round trip.
"""
import numpy as np
import pytest

from oracle import AmrTable, params, wgsl32, wgsl64
from shaderous_amd import amr_tree, problems

BS = 16
ROOT = (4, 4, 1)
DX0 = 1.0 / 64


def blob(x, y, z):
    r2 = x * x + y * y
    return {"rho": 1.0 + 0.8 * np.exp(-r2 / 0.02), "vx": 0.3 + 0.2 * np.sin(2 * np.pi * y),
            "vy": -0.2 * x + 0.1 * np.cos(2 * np.pi * x), "vz": 0.0 * x, "p": 1.0 + 2.0 * np.exp(-r2 / 0.01)}


def expanding(x, y, z):
    """The blob in a flow that diverges everywhere within |x|, |y| < 1/2 (the refined patch sits
    well inside it). At the rim of a face beside a coarser block the 3D passes see an
    unassembled edge texel as a z neighbour, whose zero pressure is an infinite jump: a shock
    where the velocity converges, which the 2D passes (no z neighbours) never see. Where the
    flow diverges no shock is flagged and the two restatements must agree."""
    r2 = x * x + y * y
    return {"rho": 1.0 + 0.8 * np.exp(-r2 / 0.02), "vx": 0.5 * x + 0.05 * np.sin(2 * np.pi * y),
            "vy": 0.5 * y + 0.05 * np.cos(2 * np.pi * x), "vz": 0.0 * x, "p": 1.0 + 2.0 * np.exp(-r2 / 0.01)}


def quadtree():
    """Levels 0..2: the root block at (1, 1) refined, one of its children again, balanced."""
    return amr_tree.build_tree(ROOT, 2, lambda L, c: (L == 0 and c == (1, 1, 0)) or (L == 1 and c == (3, 3, 0)),
                               ndim=2)


def prm2(limiter="mc", **kw):
    return params((64, 64, 1), ndim=2, dx=DX0, gamma=1.4, cfl=0.4, limiter=limiter, **kw)


def totals(blocks, u1, u2, ndim=2):
    leaf = amr_tree.leaf_mask(blocks)
    vol = np.repeat((DX0 / (1 << blocks[:, 0].astype(np.int64))) ** ndim * leaf, BS ** ndim)
    return np.concatenate([(u1.astype(np.float64) * vol[:, None]).sum(0), [(u2[:, 0].astype(np.float64) * vol).sum()]])


def test_quadtree_table():
    blocks = quadtree()
    assert set(np.unique(blocks[:, 0])) == {0, 1, 2} and not blocks[:, 3].any()
    tab = AmrTable(blocks, ROOT, ndim=2).table()
    leaf = amr_tree.leaf_mask(blocks)
    assert np.array_equal((tab[:, 4] & 2) != 0, leaf)
    assert abs(sum(4.0 ** -int(b[0]) for b in blocks[leaf]) / 16 - 1.0) < 1e-12
    for s, r in enumerate(tab):
        kids = r[12:20]
        assert np.all(kids[4:] == -1)
        assert np.all(kids[:4] >= 0) if not leaf[s] else np.all(kids == -1)
        for o in range(4):
            if kids[o] >= 0:
                L, x, y, z = blocks[kids[o]]
                assert (L, x, y, z) == (r[0] + 1, 2 * r[1] + (o & 1), 2 * r[2] + (o >> 1), 0) and tab[kids[o], 5] == s
        assert r[10] == s and r[11] == s                   # z faces: the block itself
        for f in range(4):
            n = r[6 + f]
            if n >= 0:
                assert tab[n, 6 + (f ^ 1)] == s and tab[n, 0] == r[0]
            else:
                assert n == -1 and tab[r[20 + f], 0] == r[0] - 1 and (tab[r[20 + f], 4] & 2)


def blocks_from_uniform2(u, root):
    out = []
    for by in range(root[1]):
        for bx in range(root[0]):
            out.append(u[BS * by:BS * by + BS, BS * bx:BS * bx + BS].reshape(-1, 4))
    return np.ascontiguousarray(np.concatenate(out))


@pytest.mark.parametrize("impl,tol", [(wgsl64, 1e-12), (wgsl32, 1e-5)])
@pytest.mark.parametrize("bcs", ["periodic", "walls"])
def test_unrefined_quadtree_matches_uniform(impl, tol, bcs):
    bc_lo, bc_hi = ((0, 0, 0), (0, 0, 0)) if bcs == "periodic" else ((1, 2, 0), (2, 1, 0))
    prm = prm2(bc_lo=bc_lo, bc_hi=bc_hi)
    x, y, z = problems.centers((64, 64, 1), 1.0)
    u1, u2 = problems.cons_from_prims(blob(x, y, 0.0 * z), 1.4, impl.dtype)
    b1 = blocks_from_uniform2(u1.reshape(64, 64, 4), ROOT)
    b2 = blocks_from_uniform2(u2.reshape(64, 64, 4), ROOT)
    roots = np.asarray([(0, i, j, 0) for j in range(4) for i in range(4)], np.int32)
    t = AmrTable(roots, ROOT, bc_lo, bc_hi, ndim=2)
    cu = np.zeros(4, impl.dtype)
    ca = np.zeros(4, impl.dtype)
    for _ in range(3):
        impl.step(prm, u1, u2, cu)
        impl.amr_step(prm, t, DX0, b1, b2, ca)
        assert abs(ca[0] - cu[0]) <= tol * cu[0]
        ref = blocks_from_uniform2(u1.reshape(64, 64, 4), ROOT)
        rel = np.abs(b1 - ref).max(axis=0) / np.abs(ref).max(axis=0).clip(1e-30)
        assert rel.max() <= tol, rel


def extrude(blocks, u1, u2):
    """The z extrusion of a quadtree and its state: each block (L, x, y, 0) becomes the column
    (L, x, y, z), z < 2^L, of 16^3-cell blocks holding the 16^2 plane in every layer."""
    out_b, out1, out2 = [], [], []
    for s, (L, x, y, _) in enumerate(blocks.tolist()):
        plane1 = u1[s * 256:(s + 1) * 256]
        plane2 = u2[s * 256:(s + 1) * 256]
        for z in range(1 << L):
            out_b.append((L, x, y, z))
            out1.append(np.tile(plane1, (BS, 1)))
            out2.append(np.tile(plane2, (BS, 1)))
    return np.asarray(out_b, np.int32), np.concatenate(out1), np.concatenate(out2)


@pytest.mark.parametrize("limiter", ["mc", "koren"])
def test_refined_quadtree_steps_like_its_extrusion(limiter):
    blocks = quadtree()
    u1, u2 = amr_tree.blocks_state(blocks, DX0, expanding, 1.4, np.float64, ndim=2)
    b3, e1, e2 = extrude(blocks, u1, u2)
    t2 = AmrTable(blocks, ROOT, ndim=2)
    t3 = AmrTable(b3, ROOT)
    prm = prm2(limiter)
    # the 3D passes with the 2D rate (ndim enters the step only through the signal speeds' sum)
    prm3 = params((64, 64, 16), ndim=2, dx=DX0, gamma=1.4, cfl=0.4, limiter=limiter)
    c2, c3 = np.zeros(4), np.zeros(4)
    col = {}
    for i, (L, x, y, z) in enumerate(b3.tolist()):
        col.setdefault((L, x, y), []).append(i)
    leaf = amr_tree.leaf_mask(blocks)
    for _ in range(3):
        wgsl64.amr_step(prm, t2, DX0, u1, u2, c2)
        wgsl64.amr_step(prm3, t3, DX0, e1, e2, c3)
        assert abs(c2[0] - c3[0]) <= 1e-13 * c3[0]
        for s, (L, x, y, _) in enumerate(blocks.tolist()):
            if not leaf[s]:
                continue
            want1 = u1[s * 256:(s + 1) * 256]
            want2 = u2[s * 256:(s + 1) * 256]
            for i in col[(L, x, y)]:
                g1 = e1[i * 4096:(i + 1) * 4096].reshape(BS, 256, 4)
                g2 = e2[i * 4096:(i + 1) * 4096].reshape(BS, 256, 4)
                for a, b in ((g1, want1), (g2, want2)):
                    err = np.abs(a - b[None]).max(axis=(0, 1)) / np.abs(b).max(axis=0).clip(1e-30)
                    assert err.max() < 1e-12, (s, err)


@pytest.mark.parametrize("compensated", [True, False])
def test_quadtree_conservation_fp64(compensated):
    blocks = quadtree()
    t = AmrTable(blocks, ROOT, ndim=2)
    u1, u2 = amr_tree.blocks_state(blocks, DX0, blob, 1.4, np.float64, ndim=2)
    c0 = totals(blocks, u1, u2)
    clk = np.zeros(4)
    for _ in range(4):
        wgsl64.amr_step(prm2(), t, DX0, u1, u2, clk, compensated)
    c1 = totals(blocks, u1, u2)
    scale = np.abs(c0) + np.array([0, 1, 1, 1, 0]) * c0[0]
    assert np.all(np.abs(c1 - c0) <= 1e-13 * scale), (c1 - c0) / scale


def test_quadtree_fp32_tracks_fp64():
    blocks = quadtree()
    t = AmrTable(blocks, ROOT, ndim=2)
    a1, a2 = amr_tree.blocks_state(blocks, DX0, blob, 1.4, np.float32, ndim=2)
    d1, d2 = amr_tree.blocks_state(blocks, DX0, blob, 1.4, np.float64, ndim=2)
    ca, cd = np.zeros(4, np.float32), np.zeros(4)
    for _ in range(3):
        wgsl32.amr_step(prm2(), t, DX0, a1, a2, ca)
        wgsl64.amr_step(prm2(), t, DX0, d1, d2, cd)
    leaf = amr_tree.leaf_mask(blocks).repeat(BS ** 2)
    for q in range(4):
        den = np.abs(d1[leaf, q]).sum()
        if den > 0:
            assert np.abs(a1[leaf, q] - d1[leaf, q]).sum() / den < 1e-5, q


def test_quadtree_ghosts_of_linear_field():
    """Ghost texels across refinement boundaries hold a linear density exactly: by prolongation
    from a coarser neighbour (x and y slopes only), by the mean of four children from a finer one."""
    root = (8, 8, 1)
    blocks = amr_tree.build_tree(root, 1, lambda L, c: c in ((3, 3, 0), (4, 3, 0)), ndim=2)
    t = AmrTable(blocks, root, ndim=2)
    tab = t.table()
    dx0 = 1.0 / 128
    g = np.array([0.3, -0.2])
    lin = lambda x, y, z: {"rho": 2.0 + g[0] * x + g[1] * y, "vx": 0.1 + 0 * x, "vy": -0.2 + 0 * x, "vz": 0 * x,
                           "p": 1.0 + 0 * x}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, lin, 1.4, np.float64, ndim=2)
    prm = params((128, 128, 1), ndim=2, dx=dx0, gamma=1.4, limiter="mc")
    W, _ = wgsl64.amr_assemble(prm, t, dx0, u1, u2)
    checked = {"coarse": 0, "fine": 0}
    for s, r in enumerate(tab):
        if not (r[4] & 2):
            continue
        dx = dx0 / (1 << r[0])
        for f in range(4):
            n = r[6 + f]
            kind = "coarse" if n == -1 else ("fine" if n >= 0 and not (tab[n, 4] & 2) else None)
            if kind is None:
                continue
            axis, up = f >> 1, f & 1
            for layer in (0, 1):
                idx = [slice(2, 18)] * 2
                idx[1 - axis] = 18 + layer if up else 1 - layer
                x = (np.arange(-2, 18) + 0.5 + 16 * r[1]) * dx - 0.5
                y = (np.arange(-2, 18) + 0.5 + 16 * r[2]) * dx - 0.5
                yy, xx = np.meshgrid(y, x, indexing="ij")
                want = 2.0 + g[0] * xx + g[1] * yy
                got = W[s, 0][..., 0][tuple(idx)]
                assert np.allclose(got, want[tuple(idx)], rtol=0, atol=1e-13), (s, f, layer, kind)
                assert np.allclose(W[s, 0][..., 1][tuple(idx)], 0.1, atol=1e-13)
                checked[kind] += 1
    assert checked["coarse"] > 0 and checked["fine"] > 0, checked


@pytest.mark.parametrize("impl,tol", [(wgsl64, 1e-13), (wgsl32, 2e-6)])
def test_quadtree_regrid_conserves_and_round_trips(impl, tol):
    blocks = quadtree()
    t = AmrTable(blocks, ROOT, ndim=2)
    u1, u2 = amr_tree.blocks_state(blocks, DX0, blob, 1.4, impl.dtype, ndim=2)
    prm = prm2()
    leaf = amr_tree.leaf_mask(blocks)
    up = amr_tree.adapt_tree(blocks, ROOT, np.where(leaf, 1.0, np.nan), 0.5, -1.0, 2, ndim=2)
    assert not up[:, 3].any()
    k, s, o = amr_tree.regrid_map(blocks, up)
    assert (k == 2).sum() > 0 and np.all(o[k == 2] < 4)
    f1, f2 = impl.amr_regrid(prm, t, u1, u2, k, s, o)
    c0, c1 = totals(blocks, u1, u2), totals(up, f1, f2)
    scale = np.abs(c0) + np.array([0, 1, 1, 1, 0]) * c0[0]
    assert np.all(np.abs(c1 - c0) <= tol * scale), (c1 - c0) / scale
    tu = AmrTable(up, ROOT, ndim=2)
    kb, sb, ob = amr_tree.regrid_map(up, blocks)
    assert (kb == 3).sum() > 0
    b1, b2 = impl.amr_regrid(prm, tu, f1, f2, kb, sb, ob)
    sel = leaf.repeat(BS ** 2)
    for a, b in ((b1, u1), (b2, u2)):
        err = np.abs(a[sel].astype(np.float64) - b[sel]).max(axis=0) / np.abs(b[sel]).max(axis=0).clip(1e-30)
        assert err.max() <= 4 * tol, err


def test_quadtree_regrid_prolongs_linear_fields_exactly():
    roots = np.asarray([(0, i, j, 0) for j in range(4) for i in range(4)], np.int32)
    lin = lambda x, y, z: {"rho": 2.0 + 0.5 * x + 0.25 * y, "vx": 0.0 * x, "vy": 0.0 * x, "vz": 0.0 * x,
                           "p": 1.0 + 0.0 * x}
    u1, u2 = amr_tree.blocks_state(roots, DX0, lin, 1.4, np.float64, ndim=2)
    t = AmrTable(roots, ROOT, ndim=2)
    up = amr_tree.adapt_tree(roots, ROOT, np.ones(len(roots)), 0.5, -1.0, 1, ndim=2)
    assert len(up) == 16 + 64
    k, s, o = amr_tree.regrid_map(roots, up)
    n1, _ = wgsl64.amr_regrid(prm2(), t, u1, u2, k, s, o)
    for j in np.flatnonzero(k == 2):
        x, y, _ = amr_tree.cell_centers(up[j], DX0, ndim=2)
        if np.any(np.abs(x) > 0.45) or np.any(np.abs(y) > 0.45):
            continue                                      # beside the periodic wrap
        want = 2.0 + 0.5 * x + 0.25 * y
        assert np.allclose(n1[j * 256:(j + 1) * 256, 0], want.ravel(), rtol=0, atol=1e-13)
