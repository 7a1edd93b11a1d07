"""Resuming runs on an MI355X (shaderous_amd/snapshot.py): a run resumed from a checkpoint (the
conserved state, time and step count, through a file) continues bit for bit, on the uniform grid
and on blocks; a run restarted from a snapshot of primitives (the reference's restart) follows
the uninterrupted run within the FP32 bar."""
import math

import numpy as np
import pytest

from shaderous_amd import amr_tree, problems
from shaderous_amd import snapshot as sn
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import Sim, SimConfig, config_for

pytestmark = pytest.mark.gpu

BS = 16


def uniform_pair(path):
    pb = problems.turbulence(32)
    return pb, config_for(pb, path=path, limiter="mc", isentropic=False, cfl=0.4)


def blob(x, y, z):
    r2 = x * x + y * y + z * z
    g = np.exp(-r2 / 0.02)
    return {"rho": 1.0 + 0.8 * g, "vx": 0.3 - 2.0 * x * g, "vy": -0.2 * x - 2.0 * y * g, "vz": -2.0 * z * g,
            "p": 1.0 + 2.0 * np.exp(-r2 / 0.01), "s": (x > 0).astype(float)}


def amr_setup():
    blocks = amr_tree.build_tree((2, 2, 2), 2, lambda L, c: c == (0, 0, 0) or (L == 1 and c == (1, 1, 1)))
    cfg = SimConfig(n=(32, 32, 32), ndim=3, dx=1.0 / 32, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4,
                    t_end=math.inf)
    u1, u2 = amr_tree.blocks_state(blocks, cfg.dx, blob, cfg.gamma)
    return blocks, cfg, u1, u2


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_uniform_checkpoint_resumes_bitwise(path, tmp_path):
    pb, cfg = uniform_pair(path)
    with Sim(cfg) as a:
        a.set_state(*pb.state())
        a.step(6)
        ck = sn.checkpoint(a)
        ck.save(str(tmp_path / "ck.npz"))
        a.step(6)
        want = a.read_state()
        want_clock = a.clock()
    ck = sn.load_checkpoint(str(tmp_path / "ck.npz"))
    assert ck.steps == 6
    with Sim(cfg) as b:
        sn.resume(b, ck)
        b.step(6)
        got = b.read_state()
        got_clock = b.clock()
    print(f"\n[gpu] uniform {path} resume: t {want_clock[1]} vs {got_clock[1]}", flush=True)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert got_clock[0] == want_clock[0] and got_clock[3] == want_clock[3]
    assert abs(got_clock[1] - want_clock[1]) <= 1e-6 * want_clock[1]


@pytest.mark.parametrize("path", ["fused", "passes"])
def test_amr_checkpoint_resumes_bitwise(path, tmp_path):
    blocks, cfg, u1, u2 = amr_setup()
    a = AmrSim(cfg, blocks, (2, 2, 2), path=path)
    try:
        a.set_state(u1, u2)
        a.step(5)
        ck = sn.checkpoint(a)
        ck.save(str(tmp_path / "ck.npz"))
        a.step(5)
        want = a.read_state()
        want_clock = a.clock()
    finally:
        a.destroy()
    ck = sn.load_checkpoint(str(tmp_path / "ck.npz"))
    b = AmrSim(cfg, ck.blocks, ck.root, path=path)
    try:
        sn.resume(b, ck)
        b.step(5)
        got = b.read_state()
        got_clock = b.clock()
    finally:
        b.destroy()
    leaf = amr_tree.leaf_mask(blocks).repeat(BS ** 3)
    assert np.array_equal(got[0][leaf], want[0][leaf]) and np.array_equal(got[1][leaf], want[1][leaf])
    assert got_clock[0] == want_clock[0] and got_clock[3] == want_clock[3]


def test_snapshot_restart_follows_the_run(tmp_path):
    # uniform grid
    pb, cfg = uniform_pair("fused")
    with Sim(cfg) as a:
        a.set_state(*pb.state())
        a.step(6)
        snap = sn.capture(a)
        snap.save(str(tmp_path / "u.npz"))
        a.step(6)
        want = a.read_state()
    snap = sn.load_snapshot(str(tmp_path / "u.npz"))
    assert snap.steps == 6 and snap.data["rho"].shape == (1, cfg.n[2], cfg.n[1], cfg.n[0])
    with Sim(cfg) as b:
        sn.restart(b, snap)
        b.step(6)
        got = b.read_state()
    rho_l1 = np.abs(got[0][:, 0].astype(np.float64) - want[0][:, 0]).sum() / np.abs(want[0][:, 0]).sum()
    e_l1 = np.abs(got[1][:, 0].astype(np.float64) - want[1][:, 0]).sum() / np.abs(want[1][:, 0]).sum()
    print(f"\n[gpu] uniform snapshot restart: rho L1 {rho_l1:.3e}, E L1 {e_l1:.3e}", flush=True)
    assert rho_l1 < 1e-6 and e_l1 < 1e-6
    # blocks: the leaves, in any order
    blocks, cfg, u1, u2 = amr_setup()
    a = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        a.set_state(u1, u2)
        a.step(5)
        snap = sn.capture(a)
        a.step(5)
        want = a.read_state()
    finally:
        a.destroy()
    leaf = amr_tree.leaf_mask(blocks)
    assert len(snap.levels) == int(leaf.sum()) and snap.data["press"].shape[1:] == (BS, BS, BS)
    perm = np.random.default_rng(3).permutation(len(snap.levels))
    snap.levels, snap.locations = snap.levels[perm], snap.locations[perm]
    snap.data = {k: v[perm] for k, v in snap.data.items()}
    b = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sn.restart(b, snap)
        b.step(5)
        got = b.read_state()
    finally:
        b.destroy()
    sel = leaf.repeat(BS ** 3)
    rho_l1 = np.abs(got[0][sel, 0].astype(np.float64) - want[0][sel, 0]).sum() / np.abs(want[0][sel, 0]).sum()
    print(f"[gpu] amr snapshot restart: rho L1 {rho_l1:.3e}", flush=True)
    assert rho_l1 < 1e-6
