"""The re-mesh policy in libshx (shx_amr_adapt, host code: BlockTree.adapt, src/amr/tree.ts:489-641)
against amr_tree.adapt_tree, on 3D octrees and 2D quadtrees with random indicators, periodic and
closed boundaries: the same next tree, row for row. CPU only (no GPU call)."""
import numpy as np
import pytest

from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import adapt_tree


def trees():
    blast = amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))[0]
    kh = problems.kh_adaptive(root=8, max_level=3)[0]
    return [("blast", blast, (4, 4, 4), 3, 3), ("kh", kh, (8, 8, 1), 2, 4)]


@pytest.mark.parametrize("name,blocks,root,ndim,maxlev", trees(), ids=["blast3d", "kh2d"])
@pytest.mark.parametrize("periodic", [(True, True, True), (False, True, False)])
def test_native_adapt_equals_python(name, blocks, root, ndim, maxlev, periodic):
    rng = np.random.default_rng(11)
    leaf = amr_tree.leaf_mask(blocks)
    cur = blocks
    for it in range(4):
        leaf = amr_tree.leaf_mask(cur)
        ind = np.where(leaf, rng.uniform(0, 1, len(cur)), np.nan)
        want = amr_tree.adapt_tree(cur, root, ind, 0.8, 0.3, maxlev, periodic, ndim=ndim)
        got = adapt_tree(cur, root, ind, 0.8, 0.3, maxlev, periodic, ndim=ndim)
        assert np.array_equal(got, want), (name, it, len(got), len(want))
        amr_tree.regrid_map(cur, got)                     # within one level of the current tree
        cur = got


def test_native_adapt_refuses():
    from shaderous_amd._lib import ShxError
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))
    with pytest.raises(ShxError):
        adapt_tree(blocks, (2, 2, 2), np.zeros(len(blocks)), 0.8, 0.2, 99)
    with pytest.raises(ShxError):
        adapt_tree(blocks, (2, 2, 2), np.zeros(len(blocks)), 0.8, 0.2, 2, ndim=4)
