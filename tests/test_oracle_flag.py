"""The oracle's refinement flag (amr_flag, src/amr/flag.wgsl:66-155 for gas cells: Loehner's
normalized second derivative of density and pressure on the assembled atlas) pinned by what the
estimator's definition fixes, not by a restatement of it: a field linear along an axis has no
second difference (the flag is round-off), a step between two cells has the closed form

    chi = |drho| / sqrt((|drho| + EPS (rho_b + 3 rho_a))^2 + (ndim - 1) (4 EPS rho_a)^2)

at the cell before it (rho_a beside rho_a and rho_b), the largest over the block, and a block
with no variation reports GAS_MIN = 1e-30. A wrong coefficient, a dropped axis or a dropped EPS
term changes the closed form's value. CPU only."""
import math

import numpy as np
import pytest

from shaderous_amd import amr_tree

oracle = pytest.importorskip("oracle")
from oracle import AmrTable, params, wgsl32, wgsl64  # noqa: E402

EPS = 0.01
GAMMA = 1.4


def run_flag(impl, root, ndim, fn, bc=(1, 1, 1)):
    """Flag of a single-level tree of `root` blocks (16 cells each along every axis, dx 1/16,
    the domain [0, root]/1 box) with primitives fn(x, y, z)."""
    blocks = amr_tree.build_tree(root, 0, lambda L, c: False, ndim=ndim)
    dx0 = 1.0 / 16
    u1, u2 = amr_tree.blocks_state(blocks, dx0, fn, GAMMA, impl.dtype, box_lo=0.0, ndim=ndim)
    n = tuple(16 * r for r in root[:ndim]) + ((1,) if ndim == 2 else ())
    prm = params(n, ndim=ndim, dx=dx0, gamma=GAMMA, bc_lo=bc, bc_hi=bc, limiter="mc")
    return blocks, impl.amr_flag(prm, AmrTable(blocks, root, bc, bc, ndim=ndim), dx0, u1, u2)


def uniform(rho=1.0, p=1.0):
    def fn(x, y, z):
        return {"rho": rho + 0 * x, "vx": 0 * x, "vy": 0 * x, "vz": 0 * x, "p": p + 0 * x, "s": 0 * x}
    return fn


def step_x(at, a, b, field):
    """`field` ("rho" or "p") a for x < at, b beyond; the other one 1."""
    def fn(x, y, z):
        q = np.where(x < at, a, b)
        pr = {"rho": 1.0 + 0 * x, "vx": 0 * x, "vy": 0 * x, "vz": 0 * x, "p": 1.0 + 0 * x, "s": 0 * x}
        pr[field] = q
        return pr
    return fn


def closed_form(a, b, ndim):
    d = abs(b - a)
    return d / math.sqrt((d + EPS * (b + 3 * a)) ** 2 + (ndim - 1) * (4 * EPS * a) ** 2)


@pytest.mark.parametrize("ndim", [3, 2])
def test_uniform_state_flags_gas_min(ndim):
    root = (2, 2, 2) if ndim == 3 else (2, 2, 1)
    _, f = run_flag(wgsl64, root, ndim, uniform(1.3, 0.7), bc=(0, 0, 0))
    assert np.all(f == 1e-30)


@pytest.mark.parametrize("ndim", [3, 2])
def test_linear_field_has_no_second_difference(ndim):
    # density and pressure linear in x; outflow ends copy the end cells, so only the middle
    # block (whose ghosts are its neighbours' cells) sees the field's own second difference
    root = (3, 1, 1)
    fn = lambda x, y, z: {"rho": 1.0 + 0.3 * x, "vx": 0 * x, "vy": 0 * x, "vz": 0 * x,  # noqa: E731
                          "p": 2.0 - 0.4 * x, "s": 0 * x}
    _, f = run_flag(wgsl64, root, ndim, fn)
    assert f[1] < 1e-12, f
    # the end blocks see the outflow copy: a kink, chi well above round-off
    assert f[0] > 1e-3 and f[2] > 1e-3


@pytest.mark.parametrize("ndim", [3, 2])
@pytest.mark.parametrize("field", ["rho", "p"])
def test_step_matches_the_closed_form(ndim, field):
    root = (2, 1, 1)
    a, b = 1.0, 2.5
    # the step between cells 7 and 8 of block 0 (x = 0.5 with dx 1/16): inside the block
    _, f = run_flag(wgsl64, root, ndim, step_x(0.5, a, b, field))
    want = closed_form(a, b, ndim)
    assert abs(f[0] - want) <= 1e-12 * want, (f[0], want)
    # block 1 holds b only, and its ghost beside block 0 is b too
    assert f[1] == 1e-30
    # FP32 agrees to its rounding
    _, g = run_flag(wgsl32, root, ndim, step_x(0.5, a, b, field))
    assert abs(g[0] - want) <= 1e-6 * want


def test_step_across_blocks_reads_the_ghost():
    # a step exactly on the face between the two blocks: each block sees it through its ghost
    root = (2, 1, 1)
    a, b = 1.0, 3.0
    _, f = run_flag(wgsl64, root, 3, step_x(1.0, a, b, "rho"))
    assert abs(f[0] - closed_form(a, b, 3)) <= 1e-12
    # block 1's first cell: rho_b beside rho_a (ghost) and rho_b
    d = b - a
    want1 = d / math.sqrt((d + EPS * (3 * b + a)) ** 2 + 2 * (4 * EPS * b) ** 2)
    assert abs(f[1] - want1) <= 1e-12


def test_parents_flag_zero():
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))
    dx0 = 1.0 / 32
    u1, u2 = amr_tree.blocks_state(blocks, dx0, uniform(), GAMMA, np.float64)
    prm = params((32, 32, 32), ndim=3, dx=dx0, gamma=GAMMA, limiter="mc")
    f = wgsl64.amr_flag(prm, AmrTable(blocks, (2, 2, 2)), dx0, u1, u2)
    leaf = amr_tree.leaf_mask(blocks)
    assert np.all(f[~leaf] == 0) and np.all(f[leaf] == 1e-30)
