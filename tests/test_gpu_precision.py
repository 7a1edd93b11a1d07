"""The precision statement: for every workload, the L1 difference between the production FP32 run
and an all-FP64 run of the same kernels on the GPU, per field (density, momentum, energy), with
the conserved totals of both. The Sedov 128^3 run to t = 0.05 is the precision study itself
(gamma 1.4, ambient rho 1 with 1e-4 noise, P 1e-5, E = 1 in r < 3 cells).

Each case records its numbers in precision.json in the repository's results folder when that
folder exists.
"""
import json
import os

import numpy as np
import pytest

from shaderous_amd import amr_tree, problems
from shaderous_amd.amr import AmrSim
from shaderous_amd.sim import Sim, SimConfig, config_for

pytestmark = pytest.mark.gpu

OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")


def record(name, data):
    print(name, json.dumps(data))
    if os.path.isdir(OUT):
        path = os.path.join(OUT, "precision.json")
        allv = json.load(open(path)) if os.path.exists(path) else {}
        allv[name] = data
        json.dump(allv, open(path, "w"), indent=1)


def l1(a32, a64, sel=slice(None)):
    u1a, u2a = (np.asarray(x, np.float64)[sel] for x in a32)
    u1b, u2b = (np.asarray(x, np.float64)[sel] for x in a64)
    return {"rho": float(np.abs(u1a[:, 0] - u1b[:, 0]).sum() / np.abs(u1b[:, 0]).sum()),
            "mom": float(np.abs(u1a[:, 1:4] - u1b[:, 1:4]).sum() / np.abs(u1b[:, 1:4]).sum()),
            "E": float(np.abs(u2a[:, 0] - u2b[:, 0]).sum() / np.abs(u2b[:, 0]).sum())}


def run_uniform(pb, precision, steps=None):
    cfg = config_for(pb, precision=precision, path="fused")
    dtype = np.float64 if precision == 64 else np.float32
    sim = Sim(cfg)
    try:
        sim.set_state(*pb.state(dtype))
        d0 = sim.diagnostics()
        if steps is None:
            sim.run_to_end(batch=64, max_steps=100000)
        else:
            sim.step(steps)
        d1 = sim.diagnostics()
        return sim.read_state(), d0, d1
    finally:
        sim.destroy()


def compare_uniform(name, pb, steps, tol):
    s32, a0, a1 = run_uniform(pb, 32, steps)
    s64, b0, b1 = run_uniform(pb, 64, steps)
    d = l1(s32, s64)
    data = {"l1": d, "steps32": a1.steps, "steps64": b1.steps, "t32": a1.t, "t64": b1.t,
            "mass_drift32": abs(a1.mass - a0.mass) / a0.mass, "mass_drift64": abs(b1.mass - b0.mass) / b0.mass,
            "energy_drift32": abs(a1.e_tot - a0.e_tot) / a0.e_tot, "energy_drift64": abs(b1.e_tot - b0.e_tot) / b0.e_tot}
    record(name, data)
    assert max(d.values()) <= tol, d
    assert data["mass_drift32"] < 1e-6 and data["energy_drift32"] < 1e-6
    assert data["mass_drift64"] < 1e-13 and data["energy_drift64"] < 1e-13
    return data


def test_precision_kelvin_helmholtz_256():
    compare_uniform("kelvin_helmholtz_256^2_100_steps", problems.kelvin_helmholtz(256), 100, 1e-5)


def test_precision_turbulence_256():
    compare_uniform("turbulence_256^3_50_steps", problems.turbulence(256, limiter="mc", isentropic=False, cfl=0.4),
                    50, 1e-5)


def test_precision_sedov_128():
    data = compare_uniform("sedov_128^3_t0.05", problems.sedov(128), None, 1e-5)
    assert abs(data["t32"] - 0.05) < 1e-6 and abs(data["t64"] - 0.05) < 1e-12


def test_precision_amr_blast():
    blocks, dx0, prims, _ = problems.amr_blast(seed=7)
    leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    out = {}
    for precision in (32, 64):
        dtype = np.float64 if precision == 64 else np.float32
        cfg = SimConfig(n=(128, 128, 128), ndim=3, dx=dx0, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4,
                        precision=precision)
        sim = AmrSim(cfg, blocks, (8, 8, 8))
        try:
            sim.set_state(*amr_tree.blocks_state(blocks, dx0, prims, 1.4, dtype))
            d0 = sim.diagnostics()
            sim.step(20)
            out[precision] = (sim.read_state(), d0, sim.diagnostics())
        finally:
            sim.destroy()
    d = l1(out[32][0], out[64][0], leaf)
    a0, a1 = out[32][1], out[32][2]
    data = {"l1": d, "steps": 20, "mass_drift32": abs(a1.mass - a0.mass) / a0.mass,
            "energy_drift32": abs(a1.e_tot - a0.e_tot) / a0.e_tot,
            "mass_drift64": abs(out[64][2].mass - out[64][1].mass) / out[64][1].mass}
    record("amr_blast_20_steps", data)
    assert max(d.values()) <= 1e-5, d
    assert data["mass_drift32"] < 1e-6 and data["energy_drift32"] < 1e-6 and data["mass_drift64"] < 1e-13
