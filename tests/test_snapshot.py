"""Snapshots and checkpoints on the host (shaderous_amd/snapshot.py): the primitive <-> conserved
conversions invert each other, and both file forms read back exactly without pickle. CPU only."""
import numpy as np

from shaderous_amd import snapshot as sn


def _state(n, rng):
    rho = rng.uniform(0.1, 10.0, n)
    v = rng.uniform(-2.0, 2.0, (n, 3))
    p = rng.uniform(0.01, 100.0, n)
    s = rng.uniform(0.0, 1.0, n)
    return {"rho": rho, "press": p, "vel1": v[:, 0], "vel2": v[:, 1], "vel3": v[:, 2], "r0": s}


def test_prims_and_cons_invert_each_other():
    rng = np.random.default_rng(5)
    prims = _state(1000, rng)
    for gamma in (1.4, 5.0 / 3.0):
        u1, u2 = sn.cons_of(prims, gamma)
        # E = rho e + rho |v|^2 / 2 with rho e = p / (gamma - 1)
        ke = 0.5 * prims["rho"] * (prims["vel1"] ** 2 + prims["vel2"] ** 2 + prims["vel3"] ** 2)
        assert np.allclose(u2[:, 0], prims["press"] / (gamma - 1) + ke, rtol=1e-14)
        back = sn.prims_of(u1, u2, gamma)
        for k in sn.VARIABLES:
            assert np.allclose(back[k], prims[k], rtol=1e-13, atol=1e-15), k
    # the pressure floor applies on the way back
    prims["press"][:10] = 0.0
    u1, u2 = sn.cons_of(prims, 1.4)
    assert np.all(sn.prims_of(u1, u2, 1.4, p_floor=1e-14)["press"][:10] == 1e-14)


def test_snapshot_file_reads_back_exactly(tmp_path):
    rng = np.random.default_rng(6)
    data = {k: rng.standard_normal((3, 16, 16, 16)).astype(np.float32) for k in sn.VARIABLES}
    snap = sn.Snapshot(ndim=3, root_size=(64, 64, 64), block_size=(16, 16, 16), lo=(-0.5,) * 3, hi=(0.5,) * 3,
                       time=0.125, steps=17, levels=np.array([0, 1, 1], np.int32),
                       locations=np.array([[0, 0, 0], [2, 3, 1], [3, 3, 1]], np.int32), data=data)
    path = str(tmp_path / "s.npz")
    snap.save(path)
    back = sn.load_snapshot(path)   # np.load(allow_pickle=False) inside
    assert (back.ndim, back.root_size, back.block_size, back.time, back.steps) == (3, (64, 64, 64), (16, 16, 16), 0.125, 17)
    assert back.lo == (-0.5,) * 3 and back.hi == (0.5,) * 3
    assert np.array_equal(back.levels, snap.levels) and np.array_equal(back.locations, snap.locations)
    for k in sn.VARIABLES:
        assert back.data[k].dtype == np.float32 and np.array_equal(back.data[k], data[k])


def test_checkpoint_file_reads_back_exactly(tmp_path):
    rng = np.random.default_rng(7)
    u1 = rng.standard_normal((4096 * 2, 4)).astype(np.float32)
    u2 = rng.standard_normal((4096 * 2, 4)).astype(np.float32)
    blocks = np.array([[0, 0, 0, 0], [0, 1, 0, 0]], np.int32)
    ck = sn.Checkpoint(u1=u1, u2=u2, time=0.3125, steps=40, blocks=blocks, root=(2, 1, 1))
    path = str(tmp_path / "c.npz")
    ck.save(path)
    back = sn.load_checkpoint(path)
    assert np.array_equal(back.u1, u1) and np.array_equal(back.u2, u2) and back.u1.dtype == np.float32
    assert (back.time, back.steps, back.root) == (0.3125, 40, (2, 1, 1)) and np.array_equal(back.blocks, blocks)
    flat = sn.Checkpoint(u1=u1, u2=u2, time=1.0, steps=2)
    flat.save(path)
    assert sn.load_checkpoint(path).blocks is None
