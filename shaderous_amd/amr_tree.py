"""Block trees for the AMR hydro path: 16^3-cell blocks, each refined into 8 children (3D), or
16^2-cell blocks refined into 4 (2D: iz = 0 throughout and one root block along z).

A tree is the list of its blocks as (level, ix, iy, iz), parents and leaves, in the reference's
convention (src/amr/tree.ts:14-53): block coordinates count blocks of that level from the
domain's lower corner, the root grid has `root` blocks per side, and a block of level L covers
cells [16 c 2^-L dx0 ...). Trees are built on the host and must keep 2:1 balance across faces
(tree.ts neighbor :293-324). A run re-meshes by handing the sim a new tree that differs from the
old one by at most one level everywhere (`adapt_tree`, `regrid_map`).
"""
from __future__ import annotations

import numpy as np

BS = 16


def _kids(c, ndim):
    """The children's coordinates (octant x | y << 1 | z << 2) of block coordinates c."""
    return [(2 * c[0] + (o & 1), 2 * c[1] + ((o >> 1) & 1), 2 * c[2] + ((o >> 2) & 1) if ndim == 3 else 0)
            for o in range(1 << ndim)]


def _ndim_of(blocks):
    """3, or 2 for a tree whose blocks all sit at iz = 0 with four children per parent."""
    blocks = np.asarray(blocks)
    if len(blocks) == 0 or blocks[:, 3].any():
        return 3
    keys = {tuple(b) for b in blocks.tolist()}
    for L, x, y, z in blocks.tolist():
        if (L + 1, 2 * x, 2 * y, 0) in keys:
            return 3 if (L + 1, 2 * x, 2 * y, 1) in keys else 2
    return 3


def _face_neighbors(level, c, root, periodic, ndim=3):
    out = []
    for axis in range(ndim):
        n = root[axis] << level
        for step in (-1, 1):
            v = c[axis] + step
            if v < 0 or v >= n:
                if not periodic[axis]:
                    continue
                v %= n
            d = list(c)
            d[axis] = v
            out.append(tuple(d))
    return out


def build_tree(root, max_level, refine, periodic=(True, True, True), ndim=3):
    """Refine root blocks while `refine(level, (ix, iy, iz))` says so, up to `max_level`, then
    refine further wherever 2:1 face balance requires. Returns an int32 array (nblocks, 4) of
    (level, ix, iy, iz): every block of the tree, parents first by level. ndim 2: a quadtree
    (root[2] must be 1)."""
    root = tuple(int(r) for r in root)
    if ndim == 2 and root[2] != 1:
        raise ValueError("a 2D tree has one root block along z")
    leaves = {(0, (i, j, k)) for k in range(root[2]) for j in range(root[1]) for i in range(root[0])}
    refined = set()
    for level in range(max_level):
        want = sorted(c for (L, c) in leaves if L == level and refine(level, c))
        for c in want:
            _split(leaves, refined, level, c, ndim)
    # 2:1 balance: a leaf whose face neighbour region is covered by leaves two levels finer
    changed = True
    while changed:
        changed = False
        for (L, c) in sorted(leaves):
            if L == 0:
                continue
            for d in _face_neighbors(L, c, root, periodic, ndim):
                # the region beyond this face at level L must be a block (leaf or parent) of
                # level L, or covered by a leaf of level L - 1
                if (L, d) in leaves or (L, d) in refined:
                    continue
                pc = (L - 1, (d[0] >> 1, d[1] >> 1, d[2] >> 1))
                if pc in leaves:
                    continue
                # covered by something coarser than L - 1: split that leaf
                for up in range(2, L + 1):
                    q = (L - up, (d[0] >> up, d[1] >> up, d[2] >> up))
                    if q in leaves:
                        _split(leaves, refined, q[0], q[1], ndim)
                        changed = True
                        break
            if changed:
                break
    blocks = [(L, *c) for (L, c) in refined] + [(L, *c) for (L, c) in leaves]
    blocks.sort(key=lambda b: (b[0], b[3], b[2], b[1]))
    return np.asarray(blocks, dtype=np.int32)


def _split(leaves, refined, level, c, ndim=3):
    if (level, c) not in leaves:
        return
    leaves.discard((level, c))
    refined.add((level, c))
    for k in _kids(c, ndim):
        leaves.add((level + 1, k))


def leaf_mask(blocks):
    """True for the blocks of `blocks` that have no children."""
    keys = {tuple(b) for b in blocks.tolist()}
    out = np.ones(len(blocks), bool)
    for i, (L, x, y, z) in enumerate(blocks.tolist()):
        for o in range(8):
            if (L + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1)) in keys:
                out[i] = False
                break
    return out


def cell_centers(block, dx0, box_lo=-0.5, ndim=3):
    """Centres (x, y, z), each (16, 16, 16) indexed [k, j, i], of a block's cells; in 2D each
    (1, 16, 16) with z = 0."""
    L, bx, by, bz = (int(v) for v in block)
    dx = dx0 / (1 << L)
    ax = [(np.arange(BS) + 0.5 + BS * c) * dx + box_lo for c in (bx, by, bz)]
    if ndim == 2:
        ax[2] = np.zeros(1)
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return x, y, z


def blocks_state(blocks, dx0, prim_fn, gamma, dtype=np.float32, box_lo=-0.5, ndim=3):
    """Leaf-interior state of every block (parents get zeros): (nblocks * 16^ndim, 4) u1 and u2,
    in the block order given, cells x fastest. prim_fn(x, y, z) -> dict of primitives
    (vectorised); in 2D z is 0."""
    from .problems import cons_from_prims
    n = len(blocks)
    nc = BS ** ndim
    u1 = np.zeros((n * nc, 4), dtype)
    u2 = np.zeros((n * nc, 4), dtype)
    leaf = np.flatnonzero(leaf_mask(blocks))
    ax = np.arange(BS, dtype=np.float64) + 0.5
    k, j, i = np.meshgrid(ax if ndim == 3 else np.zeros(1), ax, ax, indexing="ij")
    cell = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)          # (16^ndim, 3), x fastest
    for lo in range(0, len(leaf), 512):                                  # bounded memory
        sel = leaf[lo:lo + 512]
        b = blocks[sel].astype(np.float64)
        h = dx0 / np.exp2(b[:, 0])[:, None, None]                        # (m, 1, 1)
        pos = (cell[None, :, :] + BS * b[:, None, 1:4]) * h + box_lo     # (m, 16^ndim, 3)
        if ndim == 2:
            pos[..., 2] = 0.0
        a1, a2 = cons_from_prims(prim_fn(pos[..., 0], pos[..., 1], pos[..., 2]), gamma, dtype)
        rows = (sel[:, None] * nc + np.arange(nc)[None, :]).ravel()
        u1[rows] = a1
        u2[rows] = a2
    return u1, u2


def blast_tree(root=8, max_level=3, radius=0.1, center=(0.0, 0.0, 0.0), threshold=0.3, p_in=100.0, p_out=1.0,
               buffer=0.5):
    """The block-AMR blast workload's tree: root^3 blocks on the unit box, refined up to
    `max_level` wherever the relative jump of the initial state across a block (the pressure's:
    the initial density is uniform) exceeds `threshold` within `buffer` block widths of it: the
    blocks that straddle the sphere's surface, grown by half a block on every side. With the
    defaults this gives 5000 blocks (4439 leaves, 18.2 M leaf cells) for the seeded centre."""
    dx0 = 1.0 / (root * BS)
    c = np.asarray(center)

    def refine(level, b):
        size = BS * dx0 / (1 << level)
        lo = np.asarray(b) * size - 0.5
        # nearest and farthest distance of the block (grown by `buffer` blocks) from the centre
        glo, ghi = lo - buffer * size, lo + (1 + buffer) * size
        near = np.sqrt(np.sum(np.maximum(0.0, np.maximum(glo - c, c - ghi)) ** 2))
        far = np.sqrt(np.sum(np.maximum(np.abs(glo - c), np.abs(ghi - c)) ** 2))
        # the region straddles the sphere's surface: a pressure jump p_in/p_out there
        return near <= radius <= far and abs(p_in - p_out) / min(p_in, p_out) > threshold
    return build_tree((root, root, root), max_level, refine), dx0


def regrid_map(old_blocks, new_blocks):
    """What re-meshing does for each block of the new tree (src/amr/regrid.wgsl:40-139, the
    table of tree.ts:681-720): kind 0 (not a leaf), 1 copy of an old leaf, 2 prolongation of an
    old leaf into its child `oct`, 3 restriction of an old parent whose children are all old
    leaves. Returns int32 arrays (kind, src, oct), src an old slot. Raises ValueError for a new
    leaf more than one level from the old tree."""
    old = {tuple(b): i for i, b in enumerate(np.asarray(old_blocks).tolist())}
    old_leaf = leaf_mask(np.asarray(old_blocks))
    ndim = min(_ndim_of(old_blocks), _ndim_of(new_blocks))
    new_blocks = np.asarray(new_blocks)
    new_leaf = leaf_mask(new_blocks)
    n = len(new_blocks)
    kind = np.zeros(n, np.int32)
    src = np.zeros(n, np.int32)
    oct_ = np.zeros(n, np.int32)
    for j, (L, x, y, z) in enumerate(new_blocks.tolist()):
        if not new_leaf[j]:
            continue
        i = old.get((L, x, y, z))
        if i is not None and old_leaf[i]:
            kind[j], src[j] = 1, i
            continue
        if i is not None:
            kids = [old.get((L + 1, *k)) for k in _kids((x, y, z), ndim)]
            if all(k is not None and old_leaf[k] for k in kids):
                kind[j], src[j] = 3, i
                continue
        p = old.get((L - 1, x >> 1, y >> 1, z >> 1)) if L > 0 else None
        if p is not None and old_leaf[p]:
            kind[j], src[j], oct_[j] = 2, p, (x & 1) | ((y & 1) << 1) | ((z & 1) << 2)
            continue
        raise ValueError(f"new leaf {(L, x, y, z)} is more than one level from the old tree")
    return kind, src, oct_


def adapt_tree(blocks, root, indicator, refine_above, coarsen_below, max_level, periodic=(True, True, True),
               ndim=None):
    """The next tree from the current one and a per-block indicator (NaN for parents): leaves
    above `refine_above` split (below `max_level`), parents whose children are leaves all
    below `coarsen_below` merge, then 2:1 balance restored by splitting (which can undo a merge).
    Every block stays within one level of the current tree. ndim defaults to the tree's own."""
    blocks = np.asarray(blocks)
    if ndim is None:
        ndim = 2 if root[2] == 1 and _ndim_of(blocks) == 2 else 3
    leaf = leaf_mask(blocks)
    keys = {tuple(b): i for i, b in enumerate(blocks.tolist())}
    split = {(L, (x, y, z)) for (L, x, y, z), lf, v in zip(blocks.tolist(), leaf, indicator)
             if lf and L < max_level and v > refine_above}
    parents = {(L, (x, y, z)) for (L, x, y, z), lf in zip(blocks.tolist(), leaf) if not lf}
    merge = set()
    for (L, c) in parents:
        kids = [keys.get((L + 1, *k)) for k in _kids(c, ndim)]
        if all(k is not None and leaf[k] and indicator[k] < coarsen_below for k in kids):
            merge.add((L, c))

    def refine(level, c):
        return ((level, c) in parents and (level, c) not in merge) or (level, c) in split
    return build_tree(root, max_level, refine, periodic, ndim)
