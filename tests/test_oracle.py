"""Pins for the oracle (CPU restatements of the reference's hydro path), on the CPU.

Each check is one the reference's own tests make (tests/unit/reference.test.ts,
src/validation/tests.ts), a closed form, or an invariant, chosen so a plausible slip (a dropped
term, a wrong sign or index, a swapped operand) fails one of them. Measured reference values come
from tests/golden/reference_measured.json.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle import RefHydro, exact_riemann, params, wgsl32, wgsl64
from shaderous_amd import problems

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_measured.json")))


# ---------------- exact Riemann solver: reference.test.ts:20-35 ----------------

def test_exact_sod_star_state():
    g = GOLD["exact_riemann"]
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    star = exact_riemann(L, R, 1.4, 0.5)
    assert abs(star[2] - g["sod_p_star"]) < 0.5e-4
    assert abs(star[1] - g["sod_u_star"]) < 0.5e-4
    assert exact_riemann(L, R, 1.4, -2.0)[0] == 1.0
    assert exact_riemann(L, R, 1.4, 2.0)[0] == 0.125
    assert abs(exact_riemann(L, R, 1.4, 1.0)[0] - g["sod_rho_between_contact_and_shock"]) < 0.5e-4


def test_exact_strong_shock():
    g = GOLD["exact_riemann"]
    star = exact_riemann((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 1.4, 19.0)
    assert abs(star[2] - g["strong_p_star"]) < 0.5e-2
    assert abs(star[1] - g["strong_u_star"]) < 0.5e-3


def test_exact_rankine_hugoniot():
    """Mass and momentum fluxes are continuous across the sampled right shock (Sod)."""
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    g = 1.4
    post = exact_riemann(L, R, g, 1.7)   # between contact (0.927) and shock (1.752)
    pre = exact_riemann(L, R, g, 1.9)
    assert pre == R
    rho1, u1, p1 = post
    rho0, u0, p0 = pre
    s = (rho1 * u1 - rho0 * u0) / (rho1 - rho0)   # shock speed from mass
    mom_l = rho1 * u1 * (u1 - s) + p1
    mom_r = rho0 * u0 * (u0 - s) + p0
    assert abs(mom_l - mom_r) < 1e-10
    assert 1.70 < s < 1.80


# ---------------- limiters: reference.test.ts:38-59 ----------------

@pytest.mark.parametrize("impl", [wgsl64, wgsl32])
@pytest.mark.parametrize("lim", ["mc", "vanleer", "minmod", "koren"])
def test_limiter_invariants(impl, lim):
    e = lambda a, b, s=0.0: float(impl.edge4(lim, a, b, s)[0])
    assert e(1, -1) == 0
    assert e(0, 3) == 0
    tol = 1e-15 if impl is wgsl64 else 1e-6
    for a, b in [(1, 2), (2, 1), (-1, -4), (0.3, 0.3)]:
        s = e(a, b)
        assert math.copysign(1, s) == math.copysign(1, a) and s != 0
        assert abs(s) <= 2 * min(abs(a), abs(b)) + tol
    assert abs(e(0.3, 0.3) - 0.3) < (1e-14 if impl is wgsl64 else 1e-7)


def test_koren_smoothness_switch():
    k = lambda far, near, s: float(wgsl64.edge4("koren", far, near, s)[0])
    assert abs(k(0.001, -0.002, 1) - (0.001 - 0.004) / 3) < 1e-14
    assert k(0.001, -0.002, 0) == 0
    assert k(0.5, -0.5, 1) == 0


def test_mc_closed_form():
    """MC: min(2|a|, 2|b|, |a + b| / 2) with the sign, zero at extrema (limiters/mc.wgsl)."""
    rng = np.random.default_rng(1)
    for a, b in rng.normal(size=(200, 2)):
        want = 0.0 if a * b <= 0 else math.copysign(min(2 * abs(a), 2 * abs(b), 0.5 * abs(a + b)), a)
        assert float(wgsl64.edge4("mc", a, b, 0)[0]) == want


# ---------------- Riemann solvers: closed forms ----------------

P14 = params((4, 1, 1), gamma=1.4)


@pytest.mark.parametrize("rs", ["hllc", "hll"])
def test_riemann_supersonic_upwind(rs):
    prm = params((4, 1, 1), gamma=1.4, riemann=rs)
    l = [1.0, 5.0, 0.3, -0.2, 1.0, 0.7]     # Mach > 4 to the right
    r = [0.5, 5.5, 0.1, 0.0, 0.8, 0.1]
    F = wgsl64.riemann(prm, l, r)
    rho, u, v, w, p, s = l
    eint = p / 0.4
    E = eint + 0.5 * rho * (u * u + v * v + w * w)
    want = [rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u, rho * u * eint / rho, rho * u * s, u]
    assert np.allclose(F, want, rtol=1e-14, atol=0)


def test_hllc_stationary_contact():
    """A contact at rest: HLLC carries no mass and exactly the pressure (hllc.wgsl:60-74)."""
    prm = params((4, 1, 1), gamma=1.4)
    F = wgsl64.riemann(prm, [1.0, 0.0, 0.3, 0.0, 1.0, 1.0], [0.1, 0.0, -0.4, 0.0, 1.0, 0.0], 0.0)
    assert F[0] == 0.0 and F[4] == 0.0 and F[7] == 0.0
    assert F[1] == 1.0
    # HLL smears it: a mass flux appears
    Fh = wgsl64.riemann(prm, [1.0, 0.0, 0.3, 0.0, 1.0, 1.0], [0.1, 0.0, -0.4, 0.0, 1.0, 0.0], 1.0)
    assert abs(Fh[0]) > 1e-3


def test_hllc_blend_endpoints():
    """w = 1 gives the HLL flux (hllc.wgsl:75-77 against hll.wgsl), w = 0 the star flux; the
    blend is linear in w."""
    l = [1.0, 0.2, 0.1, 0.0, 1.0, 1.0]
    r = [0.3, -0.1, 0.0, 0.2, 0.4, 0.0]
    hllc = params((4, 1, 1), gamma=1.4, riemann="hllc")
    hll = params((4, 1, 1), gamma=1.4, riemann="hll")
    F1 = wgsl64.riemann(hllc, l, r, 1.0)
    Fh = wgsl64.riemann(hll, l, r, 0.0)
    assert np.allclose(F1[:5], Fh[:5], rtol=1e-13, atol=1e-15)
    F0 = wgsl64.riemann(hllc, l, r, 0.0)
    Fm = wgsl64.riemann(hllc, l, r, 0.25)
    assert np.allclose(Fm[:5], 0.75 * F0[:5] + 0.25 * F1[:5], rtol=1e-13, atol=1e-15)


def test_hllc_consistency_and_symmetry():
    """Equal states give the exact Euler flux; mirroring the problem negates the normal fluxes."""
    prm = params((4, 1, 1), gamma=1.4)
    st = [0.7, 0.3, -0.2, 0.1, 2.0, 0.5]
    F = wgsl64.riemann(prm, st, st)
    rho, u, v, w, p, s = st
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    assert np.allclose(F[:5], [rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u], rtol=1e-14)
    l = [1.0, 0.2, 0.1, 0.0, 1.0, 1.0]
    r = [0.3, -0.1, 0.0, 0.2, 0.4, 0.0]
    Fa = wgsl64.riemann(prm, l, r)
    ml = [r[0], -r[1], r[2], r[3], r[4], r[5]]
    mr = [l[0], -l[1], l[2], l[3], l[4], l[5]]
    Fb = wgsl64.riemann(prm, ml, mr)
    assert np.allclose(Fb[[0, 4]], -Fa[[0, 4]], rtol=1e-13)
    assert np.allclose(Fb[1], Fa[1], rtol=1e-13)


def test_hllc_vacuum_face_takes_hll():
    """Cold gas whose S_L - v rounds to zero: the vacuum branch (hllc.wgsl:39-45) returns HLL."""
    prm = params((4, 1, 1), gamma=1.4)
    l = [1.0, -11.0, 0.0, 0.0, 1e-14 * 0.6, 0.0]     # c_s ~ 1e-7 at |v| = 11
    r = [1.0, 11.0, 0.0, 0.0, 1e-14 * 0.6, 0.0]
    F = np.asarray(wgsl32.riemann(prm, l, r, 0.0), np.float64)
    assert np.all(np.isfinite(F))
    hll = params((4, 1, 1), gamma=1.4, riemann="hll")
    Fh = np.asarray(wgsl32.riemann(hll, l, r, 0.0), np.float64)
    assert np.allclose(F[:5], Fh[:5], rtol=1e-6, atol=1e-12)
    assert F[7] == 0.0


# ---------------- primitives and predictor ----------------

def test_prim_closed_form_and_floor():
    prm = params((4, 1, 1), gamma=1.4, p_floor=1e-3, rho_floor=1e-6)
    a, b = wgsl64.prim(prm, [2.0, 1.0, -0.5, 0.25], [10.0, 3.0, 0.4, 0.0])
    assert np.allclose(a, [2.0, 0.5, -0.25, 0.125]) and np.allclose(b[:3], [1.2, 0.2, math.sqrt(1.4 * 1.2 / 2.0)])
    a, b = wgsl64.prim(prm, [1e-9, 0, 0, 0], [0.0, -1.0, 0.0, 0.0])
    assert a[0] == 1e-6 and b[0] == 1e-3 and abs(b[2] - math.sqrt(1.4 * 1e-3 / 1e-6)) < 1e-12


def _stencil(q, nb):
    return np.concatenate([q] + nb)


def test_predictor_uniform_flow_is_advection():
    """Uniform state: zero slopes, the half step leaves it unchanged; a linear density profile
    advected at v moves by v dt/2 (MC is exact on linear data)."""
    prm = params((8, 8, 8), gamma=1.4, limiter="mc", dx=0.1)
    q = np.array([1.0, 0.3, 0.0, 0.0, 1.0, 0.0, math.sqrt(1.4), 0.0])
    out = wgsl64.predictor(prm, _stencil(q, [q] * 6), 0.01)
    assert np.allclose(out[:6], q[:6]) and out[6] == 0 and out[7] == 0
    xm, xp = q.copy(), q.copy()
    xm[0] -= 0.01
    xp[0] += 0.01
    out = wgsl64.predictor(prm, _stencil(q, [xm, xp, q, q, q, q]), 0.02)
    # rho_half = rho - (dt/2dx) (v drho)
    assert abs(out[0] - (1.0 - 0.5 * 0.02 / 0.1 * 0.3 * 0.01)) < 1e-15
    assert abs(out[4] - 1.0) < 1e-15


def test_predictor_shock_flag_and_positivity():
    prm = params((8, 8, 8), gamma=1.4, limiter="mc", dx=0.1)
    q = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0])
    xm = np.array([1.0, 1.0, 0.0, 0.0, 3.0, 0.0, 1.0, 0.0])     # converging, p jump 2/1 > 0.5
    xp = np.array([1.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0])
    out = wgsl64.predictor(prm, _stencil(q, [xm, xp, q, q, q, q]), 1e-4)
    assert out[6] == 1.0 and out[7] == 0.0
    # diverging flow and a dt so large the half step drives the density negative: first order,
    # state unchanged
    xm2, xp2 = q.copy(), q.copy()
    xm2[1], xp2[1] = -1.0, 1.0
    out = wgsl64.predictor(prm, _stencil(q, [xm2, xp2, q, q, q, q]), 50.0)
    assert out[7] == 1.0 and np.allclose(out[:6], q[:6])


# ---------------- whole step: the WGSL restatement ----------------

def _sod_run(impl, n=64, axis=0, ndim=3, **over):
    pb, _ = problems.sod_tube(n, axis=axis, ndim=ndim), None
    kw = dict(gamma=1.4, limiter="koren", isentropic=True, cfl=0.8, p_floor=1e-12, t_end=0.2)
    kw.update(over)
    prm = params(pb.n, ndim=ndim, dx=1.0 / n, bc_lo=pb.bc_lo, bc_hi=pb.bc_hi, **kw)
    u1, u2 = pb.state(impl.dtype)
    clk = impl.arr(np.zeros(4))
    steps = 0
    while True:
        impl.step(prm, u1, u2, clk)
        if clk[0] == 0:
            break
        steps += 1
    return u1, u2, clk, steps


def _l1_vs_exact(rho, n):
    err = 0.0
    for i in range(n):
        s = (i + 0.5 - n / 2) / n
        err += abs(rho[i] - exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, s / 0.2)[0])
    return err / n


def test_wgsl_sod_matches_reference_measurement():
    """The reference's sod-axes run (n = 64, 3D, Koren, HLLC, CFL 0.8) measured L1 = 0.00789021
    in 75 steps on its FP32 GPU path. The FP32 restatement lands on it to 1e-8 with the
    isentropic reconstruction off (that run predates isentropicRecon's default, which moves the
    L1 to 0.007837): the whole step, clock included, is pinned by the reference's own output."""
    g = GOLD["sod_axes"]
    u1, u2, clk, steps = _sod_run(wgsl32, isentropic=False)
    l1 = _l1_vs_exact(u1[:, 0].astype(np.float64), 64)
    assert steps == g["steps"]
    assert abs(l1 - g["l1"]) < 5e-8, l1
    u1d, _, _, steps64 = _sod_run(wgsl64, isentropic=False)
    assert steps64 == g["steps"]
    assert abs(_l1_vs_exact(u1d[:, 0], 64) - g["l1"]) < 2e-6


@pytest.mark.parametrize("impl", [wgsl32, wgsl64])
def test_wgsl_axis_permutation_bitwise(impl):
    x = _sod_run(impl, n=48, axis=0, limiter="mc", isentropic=False)[0][:, 0]
    y = _sod_run(impl, n=48, axis=1, limiter="mc", isentropic=False)[0][:, 0]
    z = _sod_run(impl, n=48, axis=2, limiter="mc", isentropic=False)[0][:, 0]
    assert np.array_equal(x, y) and np.array_equal(x, z)


def _blob(n=(16, 16, 4), dx=1 / 16):
    x, y, z = problems.centers_dx(n, dx)
    pr = {"rho": 1.5 + np.exp(-(x * x + y * y) / 0.02), "vx": 1.0, "vy": 0.5, "vz": 0.1 * np.sin(8 * np.pi * z),
          "p": 1 + 0.5 * np.exp(-(x * x + y * y) / 0.01)}
    return pr


@pytest.mark.parametrize("lim,isen", [("mc", False), ("koren", True)])
def test_wgsl_conservation_periodic(lim, isen):
    """reference.test.ts:93-102 for the GPU scheme: periodic 16x16x4, 40 steps."""
    n = (16, 16, 4)
    prm = params(n, dx=1 / 16, gamma=1.4, limiter=lim, isentropic=isen, p_floor=1e-12)
    u1, u2 = problems.cons_from_prims(_blob(), 1.4, np.float64)
    before = [u1[:, 0].sum(), u1[:, 1].sum(), u1[:, 2].sum(), u2[:, 0].sum()]
    wgsl64.run(prm, u1, u2, 40)
    after = [u1[:, 0].sum(), u1[:, 1].sum(), u1[:, 2].sum(), u2[:, 0].sum()]
    for b, a in zip(before, after):
        assert abs(a - b) / abs(b) < 1e-12


def test_wgsl_reflecting_symmetry_and_closure():
    """reference.test.ts:135-142 for the GPU scheme."""
    n = 32
    prm = params((n, 1, 1), ndim=1, dx=1 / n, gamma=1.4, limiter="mc", isentropic=False, bc_lo=(2, 0, 0),
                 bc_hi=(2, 0, 0), p_floor=1e-12, t_end=0.5)
    x, _, _ = problems.centers_dx((n, 1, 1), 1 / n)
    u1, u2 = problems.cons_from_prims({"rho": np.ones_like(x), "p": np.where(np.abs(x) < 0.1, 10.0, 0.1)}, 1.4,
                                      np.float64)
    m0 = u1[:, 0].sum()
    clk = np.zeros(4)
    while True:
        wgsl64.step(prm, u1, u2, clk)
        if clk[0] == 0:
            break
    assert abs(u1[:, 0].sum() - m0) / m0 < 1e-12
    assert np.allclose(u1[:16, 0], u1[::-1][:16, 0], rtol=0, atol=1e-10)


def test_wgsl_face_bit_identity():
    """flux-determinism (src/validation/tests.ts:960-978): each face's mass flux computed from
    its two cells' sides is the same bits."""
    n = (16, 16, 4)
    prm = params(n, dx=1 / 16, gamma=1.4, limiter="koren", isentropic=True)
    u1, u2 = problems.cons_from_prims(_blob(), 1.4, np.float32)
    clk = np.zeros(4, np.float32)
    wgsl32.clock(prm, clk, wgsl32.max_rate(prm, u1, u2))
    pa, pb = wgsl32.predict(prm, u1, u2, clk[0])
    fm = wgsl32.face_mass(prm, u1, u2, pa, pb).reshape(4, 16, 16, 6)
    assert np.array_equal(fm[:, :, :, 1], np.roll(fm[:, :, :, 0], -1, axis=2))
    assert np.array_equal(fm[:, :, :, 3], np.roll(fm[:, :, :, 2], -1, axis=1))
    assert np.array_equal(fm[:, :, :, 5], np.roll(fm[:, :, :, 4], -1, axis=0))


def test_wgsl_fp32_tracks_fp64():
    """The FP32 scheme stays within single-precision rounding of its FP64 twin over 20 steps
    of a smooth 3D flow (the precision study's bar, 1e-5 L1)."""
    n = (16, 16, 4)
    prm = params(n, dx=1 / 16, gamma=1.4, limiter="mc", isentropic=False)
    a1, a2 = problems.cons_from_prims(_blob(), 1.4, np.float32)
    b1, b2 = a1.astype(np.float64), a2.astype(np.float64)
    wgsl32.run(prm, a1, a2, 20)
    wgsl64.run(prm, b1, b2, 20)
    l1 = np.abs(a1[:, 0] - b1[:, 0]).sum() / np.abs(b1[:, 0]).sum()
    assert l1 < 1e-5


# ---------------- the RefHydro restatement: reference.test.ts:62-149 ----------------

def _ref(n, **over):
    kw = dict(gamma=1.4, cfl=0.8, rho_floor=1e-10, p_floor=1e-12, eta_dual=1e-3, shock_thresh=0.5, riemann="hllc",
              limiter="mc", bc_lo=(1, 0, 0), bc_hi=(1, 0, 0))
    kw.update(over)
    dims = over.pop("dims", None) or (n, 1, 1)
    kw.pop("dims", None)
    h = RefHydro(params(dims, ndim=3, dx=kw.pop("dx", 1.0 / n), **kw))
    return h


def _ref_init(h, prm_fn, dims, dx, gamma=1.4):
    x, y, z = problems.centers_dx(dims, dx)
    pr = prm_fn(x, y, z)
    u1, u2 = problems.cons_from_prims(pr, gamma, np.float64)
    h.set_cons(u1, u2)


def _tube(x, left, right):
    lt = x < 0
    return {"rho": np.where(lt, left[0], right[0]), "vx": np.where(lt, left[1], right[1]),
            "p": np.where(lt, left[2], right[2]), "s": np.where(lt, 1.0, 0.0)}


def test_ref_sod_converges():
    def err(n):
        h = _ref(n)
        _ref_init(h, lambda x, y, z: _tube(x, (1, 0, 1), (0.125, 0, 0.1)), (n, 1, 1), 1 / n)
        h.run(0.2)
        assert abs(h.t - 0.2) < 1e-12
        rho = h.U(0)
        return sum(abs(rho[i] - exact_riemann((1, 0, 1), (0.125, 0, 0.1), 1.4, ((i + 0.5) / n - 0.5) / 0.2)[0])
                   for i in range(n)) / n
    e64, e128 = err(64), err(128)
    assert e128 < 0.005 and e128 < 0.7 * e64


def test_ref_second_order_smooth():
    def err(n):
        h = _ref(n, bc_lo=(0, 0, 0), bc_hi=(0, 0, 0))
        rho0 = lambda x: 1.5 + 0.2 * np.sin(2 * np.pi * x)
        _ref_init(h, lambda x, y, z: {"rho": rho0(x), "vx": 1.0, "p": 1.0}, (n, 1, 1), 1 / n)
        h.run(1.0)
        xs = (np.arange(n) + 0.5) / n - 0.5
        return np.abs(h.U(0) - rho0(xs)).sum() / n
    assert math.log2(err(32) / err(64)) > 1.7


def test_ref_conservation_periodic():
    n = (16, 16, 4)
    h = RefHydro(params(n, ndim=3, dx=1 / 16, gamma=1.4, limiter="mc", p_floor=1e-12, cfl=0.8))
    h.set_cons(*problems.cons_from_prims(_blob(), 1.4, np.float64))
    before = [h.U(i).sum() for i in (0, 1, 2, 4)]
    for _ in range(40):
        h.step()
    for b, i in zip(before, (0, 1, 2, 4)):
        assert abs(h.U(i).sum() - b) / abs(b) < 1e-12


def test_ref_positivity_pressure_ratio_1e5():
    h = _ref(128)
    _ref_init(h, lambda x, y, z: _tube(x, (1, 0, 1000), (1, 0, 0.01)), (128, 1, 1), 1 / 128)
    h.run(0.012)
    assert np.all(h.U(0) > 0.5)
    assert np.all(0.4 * h.U(5) > 0.009)


def test_ref_axis_permutation_bitwise():
    def run(axis):
        dims = [1, 1, 1]
        dims[axis] = 48
        bc = [0, 0, 0]
        bc[axis] = 1
        h = RefHydro(params(tuple(dims), ndim=3, dx=1 / 48, gamma=1.4, limiter="mc", p_floor=1e-12, bc_lo=bc, bc_hi=bc))
        def init(x, y, z):
            s = (x, y, z)[axis]
            v = [0.0, 0.0, 0.0]
            v[axis] = 0.3
            return {"rho": np.where(s < 0, 1.0, 0.125), "vx": v[0], "vy": v[1], "vz": v[2],
                    "p": np.where(s < 0, 1.0, 0.1)}
        _ref_init(h, init, tuple(dims), 1 / 48)
        h.run(0.15)
        return h.U(0).copy()
    x = run(0)
    assert np.array_equal(run(1), x) and np.array_equal(run(2), x)


def test_ref_reflecting_symmetry():
    h = _ref(32, bc_lo=(2, 0, 0), bc_hi=(2, 0, 0))
    _ref_init(h, lambda x, y, z: {"rho": np.ones_like(x), "p": np.where(np.abs(x) < 0.1, 10.0, 0.1)}, (32, 1, 1),
              1 / 32)
    m0 = h.U(0).sum()
    h.run(0.5)
    assert abs(h.U(0).sum() - m0) / m0 < 1e-12
    rho = h.U(0)
    for i in range(16):
        assert abs(rho[i] - rho[31 - i]) < 0.5e-10


# ---------------- the two restatements against each other ----------------

@pytest.mark.parametrize("lim,isen", [("mc", False), ("koren", True), ("vanleer", False), ("minmod", False)])
def test_wgsl_fp64_agrees_with_refhydro_3d(lim, isen):
    """In 3D both rates are sum|v| + 3 c_s, and on a smooth flow the face lower bound and the
    vacuum branch never act: the two independent restatements (of the WGSL and of RefHydro)
    agree to float64 round-off, which pins the predictor, the reconstruction, HLLC's star state
    (two algebraically equal forms) and the update together."""
    n = (16, 16, 4)
    kw = dict(dx=1 / 16, gamma=1.4, limiter=lim, isentropic=isen, p_floor=1e-12, cfl=0.8)
    u1, u2 = problems.cons_from_prims(_blob(), 1.4, np.float64)
    h = RefHydro(params(n, ndim=3, **kw))
    h.set_cons(u1, u2)
    prm = params(n, ndim=3, **kw)
    a1, a2 = u1.copy(), u2.copy()
    clk = np.zeros(4)
    for _ in range(10):
        wgsl64.step(prm, a1, a2, clk)
        h.step()
    r1, r2 = h.cons()
    for c in range(4):
        assert np.allclose(a1[:, c], r1[:, c], rtol=1e-11, atol=1e-12), c
    assert np.allclose(a2[:, 0], r2[:, 0], rtol=1e-11, atol=1e-12)
    assert abs(clk[1] + clk[0] - h.t) < 1e-13


def test_wgsl_sod_diagonal_vs_refhydro_fp32():
    """gpu-vs-cpu (src/validation/tests.ts:2364-2403) for the FP32 restatement: 20 steps at
    16^3 of a diagonal Sod tube against float64 RefHydro; the reference measured a density L1
    of 7.8e-8 of the mean for its GPU at 32^3 (criterion 1e-4)."""
    n = 16
    dims = (n, n, n)
    x, y, z = problems.centers_dx(dims, 1 / n)
    d = (x + y + z) / math.sqrt(3)
    pr = {"rho": np.where(d < 0, 1.0, 0.125), "p": np.where(d < 0, 1.0, 0.1), "s": np.where(d < 0, 1.0, 0.0)}
    u1, u2 = problems.cons_from_prims(pr, 1.4, np.float64)
    kw = dict(dx=1 / n, gamma=1.4, limiter="koren", isentropic=True, p_floor=1e-12, cfl=0.8, bc_lo=(1, 1, 1),
              bc_hi=(1, 1, 1))
    h = RefHydro(params(dims, ndim=3, **kw))
    h.set_cons(u1, u2)
    a1, a2 = u1.astype(np.float32), u2.astype(np.float32)
    prm = params(dims, ndim=3, **kw)
    wgsl32.run(prm, a1, a2, 20)
    for _ in range(20):
        h.step()
    r1, r2 = h.cons()
    l1 = np.abs(a1[:, 0] - r1[:, 0]).sum() / np.abs(r1[:, 0]).sum()
    le = np.abs(a2[:, 0] - r2[:, 0]).sum() / np.abs(r2[:, 0]).sum()
    assert l1 < 1e-6 and le < 1e-6, (l1, le)
