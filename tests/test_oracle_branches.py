"""Closed-form pins of the oracle's branches that smooth, low-Mach flows never reach, on the CPU:

- the dual-energy switch (hydro/dual_energy.wgsl:21-33, oracle/hydro_wgsl.c advance): where the
  thermal energy is below eta of the total (a contact carried at Mach 50), the advected internal
  energy is kept and E rebuilt from it; where it is above (Mach 20), E - KE is kept; below the
  pressure floor's internal energy the floor's;
- the tenuous-gas limits (hydro/fluff.wgsl:8-18): below the density threshold the internal energy
  is capped at rho u_max, the momentum at rho v_max (its direction kept), E rebuilt;
- the block predictor's rim mask (amr/predict.wgsl:205-213): a ghost next to a coarser block,
  predicted in place, drops the slopes along the axes on which it sits at the block's rim, so a
  field linear across the face's transverse direction, carried by a uniform flow, is stepped
  exactly everywhere but in the leaf cells behind those ghosts (and the coarse cells across,
  through the flux correction), where the error is the dropped transverse term's.

Each expected value comes from the mathematics of the case, not from the oracle's formulas.
"""
import numpy as np
import pytest

from oracle import AmrTable, params, wgsl32, wgsl64
from shaderous_amd import amr_tree

GAMMA = 5.0 / 3.0


def strip(nx=64, ny=4):
    """Cell centres x of a periodic (nx, ny, 1) strip on the unit length, flattened row by row."""
    x = (np.arange(nx) + 0.5) / nx
    return np.tile(x, ny)


def state(rho, v, p, eint_scale=1.0, dtype=np.float64):
    """Conserved state from primitives: rho e = p / (gamma - 1) times eint_scale; E = KE + the
    unscaled internal energy (so E - KE and rho e can differ)."""
    n = len(rho)
    u1 = np.zeros((n, 4))
    u2 = np.zeros((n, 4))
    u1[:, 0] = rho
    u1[:, 1:4] = rho[:, None] * np.asarray(v)[None, :]
    ke = 0.5 * rho * np.dot(v, v)
    eint = p / (GAMMA - 1.0)
    u2[:, 0] = ke + eint
    u2[:, 1] = eint * eint_scale
    return u1.astype(dtype), u2.astype(dtype)


def step(impl, prm, u1, u2, steps=1):
    clk = np.zeros(4, impl.dtype)
    for _ in range(steps):
        impl.step(prm, u1, u2, clk)
    return clk


def contact(mach, dtype, delta=2e-3, nx=64):
    """A periodic density contact (rho 0.5 .. 1.5) at uniform pressure 1 carried at the speed that
    makes the lightest gas move at `mach`; rho e set (1 + delta) below E - KE."""
    x = strip(nx)
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * x)
    v = mach * np.sqrt(GAMMA * 1.0 / 0.5)
    u1, u2 = state(rho, np.array([v, 0.0, 0.0]), np.ones_like(rho))
    # E - KE = (1 + delta) rho e: the two energies disagree by delta (formed in float64, rounded once)
    u2[:, 0] = 0.5 * u1[:, 1] ** 2 / u1[:, 0] + (1.0 + delta) * u2[:, 1]
    u1, u2 = u1.astype(dtype), u2.astype(dtype)
    prm = params((nx, 4, 1), ndim=2, gamma=GAMMA, cfl=0.4, limiter="mc", eta_dual=1e-3)
    return prm, u1, u2


@pytest.mark.parametrize("impl,ptol", [(wgsl64, 1e-11), (wgsl32, 1e-6)])
def test_dual_energy_keeps_the_advected_internal_energy_at_mach_50(impl, ptol):
    """Mach 50: E - KE is 1/1390 of E, below eta_dual = 1e-3, so the advected rho e is kept and
    E = KE + rho e. A contact at uniform pressure and velocity carries rho e by fluxes
    m (rho e / rho)_upwind = v p / (gamma - 1), the same on every face: p stays uniform to
    round-off, and after the step E - KE equals rho e (not the (1 + delta) rho e it started as)."""
    prm, u1, u2 = contact(50.0, impl.dtype)
    p0 = (GAMMA - 1.0) * u2[:, 1].astype(np.float64)
    m0 = u1[:, 0].astype(np.float64).sum()
    step(impl, prm, u1, u2, 3)
    e1 = u1.astype(np.float64)
    p = (GAMMA - 1.0) * u2[:, 1].astype(np.float64)
    assert np.abs(p / p0 - 1.0).max() < ptol
    ke = 0.5 * (e1[:, 1:4] ** 2).sum(1) / e1[:, 0]
    eth = u2[:, 0].astype(np.float64) - ke
    # E rebuilt from rho e: they agree to the rounding of E (1390 times rho e: ~1e-13 in FP64,
    # ~1e-4 in FP32), well inside the delta = 2e-3 they started apart
    assert np.abs(eth / u2[:, 1] - 1.0).max() < (1e-9 if impl is wgsl64 else 5e-4)
    assert abs(e1[:, 0].sum() - m0) <= (1e-6 if impl is wgsl32 else 1e-13) * m0


def test_dual_energy_keeps_e_minus_ke_at_mach_20():
    """Mach 20: E - KE is 1/223 of E, above eta_dual: the conservative E stays as the update made
    it (total E conserved to round-off) and rho e is overwritten by E - KE, so it takes the
    (1 + delta) offset it started with."""
    delta = 2e-3
    prm, u1, u2 = contact(20.0, np.float64, delta)
    ei0 = u2[:, 1].copy()
    e0 = u2[:, 0].sum()
    step(wgsl64, prm, u1, u2, 1)
    assert abs(u2[:, 0].sum() - e0) <= 1e-13 * e0
    ke = 0.5 * (u1[:, 1:4] ** 2).sum(1) / u1[:, 0]
    assert np.array_equal(u2[:, 1], u2[:, 0] - ke)
    # the contact keeps p uniform: rho e after is (1 + delta) times the advected one
    assert np.abs(u2[:, 1] / ei0 - (1.0 + delta)).max() < 1e-8


def test_dual_energy_floor():
    """E - KE below the pressure floor's internal energy (p_floor / (gamma - 1)) takes the
    fallback too, and an advected rho e below it is raised to it: a uniform moving gas whose
    pressure is under the floor ends with rho e = p_floor / (gamma - 1) and E = KE + that."""
    n = 32
    rho = np.full(4 * n, 2.0)
    u1, u2 = state(rho, np.array([0.7, -0.2, 0.1]), np.full(4 * n, 0.01))
    prm = params((n, 4, 1), ndim=2, gamma=GAMMA, cfl=0.4, limiter="mc", p_floor=0.05)
    step(wgsl64, prm, u1, u2, 1)
    emin = 0.05 / (GAMMA - 1.0)
    assert np.allclose(u2[:, 1], emin, rtol=1e-15, atol=0)
    ke = 0.5 * (u1[:, 1:4] ** 2).sum(1) / u1[:, 0]
    assert np.allclose(u2[:, 0], ke + emin, rtol=1e-15, atol=0)
    # density and momentum of a uniform gas do not change
    assert np.allclose(u1[:, 0], 2.0, rtol=1e-15) and np.allclose(u1[:, 1], 1.4, rtol=1e-15)


@pytest.mark.parametrize("impl,tol", [(wgsl64, 1e-14), (wgsl32, 1e-6)])
def test_fluff_limits(impl, tol):
    """fluff = (rho_t, u_max, v_max): a uniform gas (no net fluxes) below rho_t ends with
    rho e = rho u_max where it exceeded it, |m| = rho v_max in the same direction where it
    exceeded it, and E = |m|^2 / (2 rho) + rho e; gas at or above rho_t is left as it was."""
    n = 32
    fl = (1.0, 2.0, 1.0)
    prm = params((n, 4, 1), ndim=2, gamma=GAMMA, cfl=0.4, limiter="mc", fluff=fl)
    v = np.array([3.0, -4.0, 0.0])                   # |v| = 5 > v_max
    # tenuous, hot and fast: both limits
    u1, u2 = state(np.full(4 * n, 0.1), v, np.full(4 * n, 1.0), dtype=impl.dtype)
    step(impl, prm, u1, u2, 1)
    m = u1[:, 1:4].astype(np.float64)
    assert np.allclose(u2[:, 1], 0.1 * fl[1], rtol=tol, atol=0)
    assert np.allclose(np.sqrt((m ** 2).sum(1)), 0.1 * fl[2], rtol=tol, atol=0)
    assert np.allclose(m / np.sqrt((m ** 2).sum(1))[:, None], v / 5.0, rtol=0, atol=tol)
    assert np.allclose(u2[:, 0], 0.5 * (0.1 * fl[2]) ** 2 / 0.1 + 0.1 * fl[1], rtol=tol, atol=0)
    # tenuous, cold and fast: only the momentum; rho e (0.015 < 0.2) kept
    u1, u2 = state(np.full(4 * n, 0.1), v, np.full(4 * n, 0.01), dtype=impl.dtype)
    ei0 = u2[:, 1].copy()
    step(impl, prm, u1, u2, 1)
    # rho e is E - KE here (E - KE is 1.2% of E, above eta): E's rounding, 84 times rho e
    assert np.allclose(u2[:, 1], ei0, rtol=100 * tol, atol=0)
    assert np.allclose(np.sqrt((u1[:, 1:4].astype(np.float64) ** 2).sum(1)), 0.1 * fl[2], rtol=tol)
    assert np.allclose(u2[:, 0], 0.5 * 0.01 / 0.1 + ei0.astype(np.float64), rtol=tol)
    # dense: untouched (a uniform state is a fixed point of the step)
    u1, u2 = state(np.full(4 * n, 2.0), v, np.full(4 * n, 1.0), dtype=impl.dtype)
    a1, a2 = u1.copy(), u2.copy()
    step(impl, prm, u1, u2, 1)
    assert np.allclose(u1, a1, rtol=tol, atol=0) and np.allclose(u2[:, :2], a2[:, :2], rtol=tol, atol=0)


# ---------------- the block predictor's rim mask ----------------
#
# A ghost next to a coarser block (predict.wgsl:199-213) is predicted in place with the slopes
# along its rim axes (transverse axes on which it sits at 0 or 15) masked out, because its
# neighbours along them are edge texels, which the assembly leaves empty (kind 2, grid.wgsl:180-
# 183). The shock detector (predictor.wgsl:114-134) still reads those empty texels: their
# velocity is 0 and their pressure 0, so the relative pressure jump is infinite and the flag is
# set wherever the velocities it sees converge, conv = sum over the rim axes of +v_t at rim 0 and
# -v_t at rim 15 (the face's own axis adds nothing in a uniform flow). A flagged ghost makes its
# face an HLL face (flux.wgsl:93). In a uniform flow v over a density linear in y, rho = 2 + b y,
# the step is then exact but at the faces of such ghosts, where
#   - the ghost's half-step density lacks the dropped term: delta = dt v_y b / 2 if y is a rim axis;
#   - the face flux is HLL (flagged) or the upwind density times the normal velocity (HLLC on a
#     contact) of the two face densities (the exact one on the cell's side, the exact one plus
#     delta on the ghost's), against the exact flux u_n rho_face;
#   - the leaf cell takes dt/h_fine of the difference, the coarse cell across dt/h_coarse of a
#     quarter of it (the mean of the four fine fluxes under its face).

RIM_ROOT = (4, 4, 4)
RIM_DX0 = 1.0 / 64


def rim_case(v, g):
    """A 4^3 root grid with root block (1, 1, 1) refined; rho = 2 + g . x, velocity v, p = 1."""
    blocks = amr_tree.build_tree(RIM_ROOT, 1, lambda L, c: c == (1, 1, 1))
    lin = lambda x, y, z: {"rho": 2.0 + g[0] * x + g[1] * y + g[2] * z, "vx": v[0], "vy": v[1], "vz": v[2],
                           "p": 1.0}
    return blocks, lin


def hll_mass(rl, rr, u):
    """HLL mass flux (Harten, Lax, van Leer; Davis wave speeds) across a face of normal velocity
    u and pressure 1 between face densities rl and rr."""
    cl, cr = np.sqrt(GAMMA / rl), np.sqrt(GAMMA / rr)
    sl, sr = min(u - cl, u - cr), max(u + cl, u + cr)
    if sl >= 0:
        return rl * u
    if sr <= 0:
        return rr * u
    return (sr * rl * u - sl * rr * u + sl * sr * (rr - rl)) / (sr - sl)


def rim_expected(blocks, dt, v, g):
    """The density error of every leaf cell after one step, as (nblocks, 4096), from the model
    above (the dropped term is dt/2 times the sum over the rim axes of v_t g_t)."""
    keys = {tuple(r): i for i, r in enumerate(blocks.tolist())}
    leaf = amr_tree.leaf_mask(blocks)
    err = np.zeros((len(blocks), 16 ** 3))
    hf, hc = RIM_DX0 / 2, RIM_DX0
    for s, (L, bx, by, bz) in enumerate(blocks.tolist()):
        if L != 1 or not leaf[s]:
            continue
        c = (bx, by, bz)
        for n in range(3):
            for side in (-1, 1):
                d = list(c)
                d[n] += side
                if (1, *d) in keys:
                    continue
                cs = keys[(0, d[0] >> 1, d[1] >> 1, d[2] >> 1)]
                assert leaf[cs]
                ta, tb = [a for a in range(3) if a != n]
                for t1 in range(16):
                    for t2 in range(16):
                        l = [0, 0, 0]
                        l[n] = -1 if side < 0 else 16
                        l[ta], l[tb] = t1, t2
                        rim = [a for a in (ta, tb) if l[a] in (0, 15)]
                        conv = sum(v[a] if l[a] == 0 else -v[a] for a in rim)
                        shock = bool(rim) and conv < 0
                        delta = 0.5 * dt * sum(v[a] * g[a] for a in rim)
                        lc = list(l)
                        lc[n] = 0 if side < 0 else 15
                        # the face's centre and the exact face density half a step on
                        xf = [(16 * c[a] + lc[a] + 0.5) * hf - 0.5 for a in range(3)]
                        xf[n] += side * 0.5 * hf
                        rho_f = 2.0 + sum(g[a] * (xf[a] - 0.5 * dt * v[a]) for a in range(3))
                        u = v[n]
                        rl, rr = (rho_f + delta, rho_f) if side < 0 else (rho_f, rho_f + delta)
                        if shock:
                            F = hll_mass(rl, rr, u)
                        else:
                            F = u * rl if u > 0 else (u * rr if u < 0 else 0.0)
                        dF = F - u * rho_f
                        if dF == 0.0:
                            continue
                        err[s, (lc[2] * 16 + lc[1]) * 16 + lc[0]] += -side * dt / hf * dF
                        gi = [16 * c[a] + lc[a] for a in range(3)]
                        gi[n] += side
                        gc = [gi[a] // 2 - 16 * (d[a] >> 1) for a in range(3)]
                        err[cs, (gc[2] * 16 + gc[1]) * 16 + gc[0]] += side * dt / hc * dF / 4
    return err


# Gradients of both signs along each transverse axis: an unmasked slope toward an empty edge
# texel (density 0) is limited to zero unless the density falls toward it, so each sign tests
# the mask at one of an axis's two rims.
@pytest.mark.parametrize("g", [(0.0, 2.0, 0.0), (0.0, -1.5, 0.0), (0.8, -1.2, 1.5), (-0.6, 1.1, -1.4)])
@pytest.mark.parametrize("v", [(1.0, 1.0, 0.0), (1.0, 0.8, 0.0), (0.0, 1.0, 0.0), (1.0, -1.0, 0.0), (-0.5, 1.0, 0.7)])
def test_amr_rim_mask_closed_form(v, g):
    blocks, lin = rim_case(v, g)
    t = AmrTable(blocks, RIM_ROOT)
    u1, u2 = amr_tree.blocks_state(blocks, RIM_DX0, lin, GAMMA, np.float64)
    prm = params((64, 64, 64), ndim=3, gamma=GAMMA, cfl=0.4, limiter="mc")
    clk = np.zeros(4)
    wgsl64.amr_step(prm, t, RIM_DX0, u1, u2, clk)
    dt = clk[0]
    assert dt > 0
    # the exact solution: the linear field carried by v dt
    want1, _ = amr_tree.blocks_state(blocks, RIM_DX0, lambda x, y, z: lin(x - v[0] * dt, y - v[1] * dt, z - v[2] * dt),
                                     GAMMA, np.float64)
    err = (u1[:, 0] - want1[:, 0]).reshape(len(blocks), 16 ** 3)
    # away from the domain's faces (the periodic wrap breaks linearity): the root blocks inside
    leaf = amr_tree.leaf_mask(blocks)
    inner = np.array([leaf[s] and all(0 < (c >> L) < 3 for c in (x, y, z))
                      for s, (L, x, y, z) in enumerate(blocks.tolist())])
    exp = rim_expected(blocks, dt, v, g)
    scale = np.abs(exp).max()
    assert scale > 1e-6 and np.count_nonzero(exp[inner]) > 100
    # the errors where expected, to their round-off, and nothing elsewhere
    dif = np.abs(err - exp)[inner]
    assert dif.max() < 1e-9 * scale, (dif.max(), scale, np.argwhere(np.abs(err - exp) * inner[:, None] > 1e-9 * scale)[:5])
