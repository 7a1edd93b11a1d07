"""The diagnostics reduction (shx_diagnostics, shx_amr_diagnostics: the 24 channels of
hydro/diagnostics.wgsl:13-98 and amr/diagnostics.wgsl:13-117, volume-weighted, FP64 sums) against
independent host FP64 sums of the state read back (tests/hostdiag.py), on the uniform grid (2D,
3D with closed faces, the 256^3 turbulence box) and on blocks; and the conserved totals of runs,
taken from those host sums, to round-off."""
import math

import numpy as np
import pytest

import hostdiag
from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import Sim, SimConfig, config_for

pytestmark = pytest.mark.gpu


def blob3d(n=32):
    dims = (n, n, n)
    x, y, z = problems.centers(dims, 1.0)
    r2 = x * x + y * y + z * z
    g = np.exp(-r2 / 0.02)
    pr = {"rho": 1.0 + 0.5 * g, "vx": 0.3 * np.sin(2 * np.pi * y) - 2.0 * x * g,
          "vy": 0.2 * np.cos(2 * np.pi * z) - 2.0 * y * g, "vz": -0.25 * np.sin(2 * np.pi * x) - 2.0 * z * g,
          "p": 1.0 + 4.0 * np.exp(-r2 / 0.005), "s": (x > 0).astype(float)}
    return problems.Problem(id="blob", n=dims, box=1.0, gamma=5.0 / 3.0, prims=pr, limiter="mc", isentropic=False,
                            cfl=0.4)


def check_uniform(sim, cfg, rtol=1e-12):
    g = sim.diagnostics().raw
    u1, u2 = sim.read_state()
    host = hostdiag.uniform(cfg, u1, u2)
    bad = hostdiag.compare(g, host, rtol)
    assert not bad, bad
    return host[0]


@pytest.mark.parametrize("case", ["kh2d", "blob_walls", "blob_floored"])
def test_uniform_diagnostics_equal_host_sums(case):
    if case == "kh2d":
        pb = problems.kelvin_helmholtz(64)
        cfg = config_for(pb)
    else:
        pb = blob3d(32)
        cfg = config_for(pb, bc_lo=(1, 2, 1), bc_hi=(2, 1, 2))
    u1, u2 = pb.state(np.float32)
    if case == "blob_floored":
        # cells at the density floor and at the pressure floor: the floored count
        cfg = config_for(pb, rho_floor=1e-3, p_floor=1e-2)
        u1[::97, 0] = 1e-3
        u2[5::89, 1] = 0.0
    with Sim(cfg) as sim:
        sim.set_state(u1, u2)
        r0 = check_uniform(sim, cfg)
        if case == "blob_floored":
            assert r0[8] > 0
        sim.step(3)
        check_uniform(sim, cfg)
    # the weighting: mass is the sum of rho dx^ndim
    assert abs(r0[0] - (np.maximum(u1[:, 0].astype(np.float64), cfg.rho_floor) * cfg.dx ** cfg.ndim).sum()) \
        <= 1e-13 * r0[0]


def test_turbulence_256_diagnostics_and_conservation():
    """At the metric's grid (turbulence 256^3, periodic): the reduction equals the host sums, and
    mass, momentum and total energy of the host sums are conserved over 10 steps to FP32
    round-off."""
    pb = problems.turbulence(256)
    cfg = config_for(pb, limiter="mc", isentropic=False, cfl=0.4)
    u1, u2 = pb.state(np.float32)
    with Sim(cfg) as sim:
        sim.set_state(u1, u2)
        h0 = check_uniform(sim, cfg, rtol=1e-11)
        sim.step(10)
        h1 = check_uniform(sim, cfg, rtol=1e-11)
    e0, e1 = h0[4] + h0[5], h1[4] + h1[5]
    assert abs(h1[0] - h0[0]) / h0[0] < 2e-7
    assert abs(e1 - e0) / e0 < 2e-7
    for a in range(3):
        assert abs(h1[1 + a] - h0[1 + a]) < 2e-7 * h0[0] * math.sqrt(5.0 / 3.0)


def test_amr_diagnostics_equal_host_sums_and_conserve():
    blocks = amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))[0]
    n = (64, 64, 64)
    cfg = SimConfig(n=n, ndim=3, dx=1.0 / 64, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4)

    def blob(x, y, z):
        r2 = x * x + y * y + z * z
        g = np.exp(-r2 / 0.02)
        return {"rho": 1.0 + 0.8 * g, "vx": 0.3 - 2.0 * x * g, "vy": -0.2 * x - 2.0 * y * g, "vz": -2.0 * z * g,
                "p": 1.0 + 2.0 * np.exp(-r2 / 0.01), "s": (x > 0).astype(float)}
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    leaf = amr_tree.leaf_mask(blocks)
    sim = AmrSim(cfg, blocks, (4, 4, 4))
    try:
        sim.set_state(u1, u2)
        out = []
        for steps in (0, 10):
            sim.step(steps)
            g = sim.diagnostics().raw
            a1, a2 = sim.read_state()
            host = hostdiag.blocks(sim.cfg, blocks, leaf, a1, a2)
            bad = hostdiag.compare(g, host)
            assert not bad, bad
            out.append(host[0])
    finally:
        sim.destroy()
    h0, h1 = out
    assert abs(h1[0] - h0[0]) / h0[0] < 2e-7
    assert abs((h1[4] + h1[5]) - (h0[4] + h0[5])) / (h0[4] + h0[5]) < 2e-7
    # the blocks' mass is the uniform grid's for the same gas: sum of rho h^3 over leaves
    assert abs(h0[0] - (1.0 + 0.8 * (0.02 * math.pi) ** 1.5)) < 2e-3


def test_run_to_end_after_steps_counts_from_the_start():
    """runToEnd returns the clock's banked step count (src/sim.ts:3705-3710), also when entered
    after steps taken one by one; with no finite t_end it refuses."""
    pb = problems.sod_tube(64, axis=0)
    cfg = config_for(pb, t_end=0.1)
    u1, u2 = pb.state(np.float32)
    with Sim(cfg) as sim:
        sim.set_state(u1, u2)
        whole = sim.run_to_end()
        end0 = sim.read_state()
        sim.set_state(u1, u2)
        assert sim.step(7) == 7
        assert sim.run_to_end(batch=4) == whole
        end1 = sim.read_state()
    assert np.array_equal(end0[0], end1[0]) and np.array_equal(end0[1], end1[1])
    from shaderous_amd._lib import ShxError
    with Sim(config_for(pb, t_end=math.inf)) as sim:
        sim.set_state(u1, u2)
        with pytest.raises(ShxError):
            sim.run_to_end()


def test_write_params_and_build_hydro():
    """writeParams takes a new CFL from the next step on; buildHydro switches the scheme; both
    runs equal sims created with the new settings, bit for bit."""
    pb = problems.kelvin_helmholtz(64)
    cfg = config_for(pb)
    u1, u2 = pb.state(np.float32)
    cfg2 = config_for(pb, cfl=0.3)
    cfg3 = config_for(pb, cfl=0.3, limiter="koren", isentropic=True)
    with Sim(cfg) as a, Sim(cfg2) as b, Sim(cfg3) as c:
        for s in (a, b, c):
            s.set_state(u1, u2)
        a.write_params(cfg2)
        a.step(3)
        b.step(3)
        assert all(np.array_equal(x, y) for x, y in zip(a.read_state(), b.read_state()))
        a.build_hydro(cfg3)
        a.step(2)
        c.set_state(*b.read_state(), time=b.clock()[1], steps=int(b.clock()[3]))
        c.step(2)
        assert all(np.array_equal(x, y) for x, y in zip(a.read_state(), c.read_state()))
        from shaderous_amd._lib import ShxError
        with pytest.raises(ShxError):
            a.write_params(config_for(pb, n=(32, 32, 1)))


def test_amr_write_params_build_hydro_and_run_to_end():
    """On blocks: writeParams takes a new CFL from the next step on, buildHydro switches the
    scheme, each run equal to a block sim created with those settings bit for bit; runToEnd then
    stops with the clock held at t_end and returns its banked step count."""
    from dataclasses import replace
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))

    def ic(x, y, z):
        return {"rho": 1.0 + 0.5 * np.exp(-(x * x + y * y + z * z) / 0.02), "vx": 0.2, "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, ic, 1.4)
    base = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4)
    c2 = replace(base, cfl=0.3, t_end=0.02)
    c3 = replace(c2, limiter="koren", isentropic=True)
    leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    sims = [AmrSim(c, blocks, (2, 2, 2)) for c in (base, c2, c3)]
    a, b, c = sims
    try:
        a.set_state(u1, u2)
        b.set_state(u1, u2)
        a.write_params(c2)
        a.step(3)
        b.step(3)
        (a1, a2), (b1, b2) = a.read_state(), b.read_state()
        assert np.array_equal(a1[leaf], b1[leaf]) and np.array_equal(a2[leaf], b2[leaf])
        a.build_hydro(c3)
        a.step(2)
        ck = b.clock()
        c.set_state(b1, b2, time=ck[1], steps=int(ck[3]))
        c.step(2)
        (a1, a2), (g1, g2) = a.read_state(), c.read_state()
        assert np.array_equal(a1[leaf], g1[leaf]) and np.array_equal(a2[leaf], g2[leaf])
        steps = a.run_to_end()
        end = a.clock()
        assert end[0] == 0 and abs(end[1] - 0.02) < 1e-6 and steps == int(end[3])
        from shaderous_amd._lib import ShxError
        with pytest.raises(ShxError):
            a.write_params(replace(c2, precision=64))
    finally:
        for s in sims:
            s.destroy()
