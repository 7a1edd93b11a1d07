"""The Node binding (shaderous_amd/node: the N-API addon shx.node and the JS facade of the
reference's simulation API). CPU: the addon loads, exports every entry the facade calls, and the
facade's configFor keeps the reference's defaults. GPU: a Node caller steps the uniform grid and
the block grid, and its results equal the Python (ctypes) caller's bit for bit and the oracle's to
the FP32 bar."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from shaderous_amd import amr_tree, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "shaderous_amd", "node", "shx.node")
pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the built addon absent")

EXPORTS = ["version", "create", "destroy", "setState", "readState", "readStateAsync", "step", "writeParams",
           "buildHydro", "sync", "runToEnd",
           "readClock", "diagnostics", "diagnosticsAsync", "fluxMismatch", "profile", "amrCreate", "amrTable",
           "amrSetState", "amrReadState", "amrStep", "amrSync", "amrReadClock", "amrDiagnostics", "amrProfile",
           "amrStats", "amrIndicator", "amrRegrid"]


def node(code):
    r = subprocess.run([NODE, "-e", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_addon_exports_and_defaults():
    out = json.loads(node("const s=require('./shaderous_amd/node');"
                          "console.log(JSON.stringify({keys:Object.keys(s.native),v:s.version(),"
                          "cfg:s.configFor(64,{ndim:2})}))"))
    assert sorted(out["keys"]) == sorted(EXPORTS)
    assert out["v"].startswith("shx ")
    c = out["cfg"]
    assert c["limiter"] == "koren" and c["riemann"] == "hllc" and c["isentropicRecon"] is True
    assert c["cfl"] == 0.8 and c["etaDual"] == 1e-3 and c["shockThresh"] == 0.5 and c["eosParams"]["gamma"] == 1.4


def test_facade_serves_the_simulation_members():
    """Every member of the reference's Simulation (src/sim.ts:465-627) that its run loops call
    on this path (src/ui/api.ts:669-699, src/validation/tests.ts withSim) is on both facades,
    and the ones the reference makes asynchronous return promises."""
    out = json.loads(node(
        "const s=require('./shaderous_amd/node');"
        "const names=['encodeSteps','submitted','stepsResume','readDiagnostics','diagnostics','reset',"
        "'readState','setState','step','destroy'];"
        "const u=names.filter(n=>typeof s.Sim.prototype[n]!=='function');"
        "const a=names.filter(n=>typeof s.AmrSim.prototype[n]!=='function');"
        "const only=['runToEnd','writeParams','buildHydro','fluxMismatch','profile']"
        ".filter(n=>typeof s.Sim.prototype[n]!=='function');"
        "const isAsync=f=>f.constructor.name==='AsyncFunction';"
        "const asy=['readState','diagnostics','readDiagnostics','runToEnd','buildHydro','fluxMismatch','profile']"
        ".filter(n=>!isAsync(s.Sim.prototype[n]));"
        "console.log(JSON.stringify({u,a,only,asy}))"))
    assert out == {"u": [], "a": [], "only": [], "asy": []}, out


def test_addon_rejects_bad_arguments():
    out = node("const s=require('./shaderous_amd/node'); const r=[];"
               "try{s.native.step({},1)}catch(e){r.push(e.constructor.name)}"
               "try{s.configFor(8,{limiter:'nope'}); new s.Sim(s.configFor(8,{limiter:'nope'}))}"
               "catch(e){r.push(e.message)}"
               "console.log(JSON.stringify(r))")
    r = json.loads(out)
    assert r[0] == "TypeError" and "unknown limiter" in r[1]


def _run(tmp_path, job, u1, u2):
    d = tmp_path / "job"
    d.mkdir()
    (d / "job.json").write_text(json.dumps(job))
    u1.tofile(d / "u1.bin")
    u2.tofile(d / "u2.bin")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "steps.js"), str(d)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    dt = u1.dtype
    o1 = np.fromfile(d / "o1.bin", dt).reshape(-1, 4)
    o2 = np.fromfile(d / "o2.bin", dt).reshape(-1, 4)
    return o1, o2, json.loads((d / "out.json").read_text())


@pytest.mark.gpu
def test_node_uniform_matches_python_and_oracle(tmp_path):
    import oracle
    from shaderous_amd.sim import Sim, config_for
    pb = problems.kelvin_helmholtz(64)
    u1, u2 = pb.state(np.float32)
    job = {"kind": "uniform", "steps": 3,
           "cfg": {"n": 64, "ndim": 2, "bcLo": list(pb.bc_lo), "bcHi": list(pb.bc_hi), "limiter": pb.limiter,
                   "isentropicRecon": pb.isentropic, "eosParams": {"gamma": pb.gamma}, "cfl": pb.cfl,
                   "box": pb.box}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    sim = Sim(config_for(pb))
    try:
        sim.set_state(u1, u2)
        sim.step(3)
        p1, p2 = sim.read_state()
    finally:
        sim.destroy()
    assert np.array_equal(o1, p1) and np.array_equal(o2, p2)
    assert out["flux"]["mismatched"] == 0 and out["flux"]["faces"] > 0
    prm = oracle.params(pb.n, ndim=2, dx=pb.dx, gamma=pb.gamma, cfl=pb.cfl, limiter=pb.limiter,
                        isentropic=pb.isentropic, bc_lo=pb.bc_lo, bc_hi=pb.bc_hi)
    r1, r2 = u1.copy(), u2.copy()
    oracle.wgsl32.run(prm, r1, r2, 3)
    l1 = np.abs(o1[:, 0] - r1[:, 0]).sum() / np.abs(r1[:, 0]).sum()
    assert l1 < 1e-5
    assert abs(out["d1"]["mass"] - out["d0"]["mass"]) / out["d0"]["mass"] < 1e-6


@pytest.mark.gpu
def test_node_amr_matches_python(tmp_path):
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))

    def ic(x, y, z):
        return {"rho": 1.0 + 0.5 * np.exp(-(x * x + y * y + z * z) / 0.02), "vx": 0.2, "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, ic, 1.4)
    job = {"kind": "amr", "steps": 4, "cfg": {"n": 32, "ndim": 3, "limiter": "mc", "isentropicRecon": False,
                                              "cfl": 0.4, "box": 1.0},
           "tree": {"root": [2, 2, 2], "blocks": blocks.reshape(-1).tolist()}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        sim.step(4)
        p1, p2 = sim.read_state()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    assert np.array_equal(o1[leaf], p1[leaf]) and np.array_equal(o2[leaf], p2[leaf])


@pytest.mark.gpu
def test_node_amr_regrid_matches_python(tmp_path):
    """A re-mesh through the Node facade (amr.regrid) then two steps on the new tree give the
    ctypes caller's bits."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))
    new = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c in ((0, 0, 0), (1, 0, 0)))

    def ic(x, y, z):
        return {"rho": 1.0 + 0.5 * np.exp(-(x * x + y * y + z * z) / 0.02), "vx": 0.2, "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, ic, 1.4)
    job = {"kind": "amr", "steps": 2, "after": 2, "regrid": new.reshape(-1).tolist(),
           "cfg": {"n": 32, "ndim": 3, "limiter": "mc", "isentropicRecon": False, "cfl": 0.4, "box": 1.0},
           "tree": {"root": [2, 2, 2], "blocks": blocks.reshape(-1).tolist(), "path": "fused"}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        sim.step(2)
        sim.regrid(new)
        sim.step(2)
        p1, p2 = sim.read_state()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(new).repeat(4096)
    assert o1.shape == p1.shape
    assert np.array_equal(o1[leaf], p1[leaf]) and np.array_equal(o2[leaf], p2[leaf])


def _kh_job(pb, **extra):
    job = {"kind": "uniform", "steps": 3,
           "cfg": {"n": pb.n[0], "ndim": 2, "bcLo": list(pb.bc_lo), "bcHi": list(pb.bc_hi), "limiter": pb.limiter,
                   "isentropicRecon": pb.isentropic, "eosParams": {"gamma": pb.gamma}, "cfl": pb.cfl,
                   "box": pb.box}}
    job.update(extra)
    return job


@pytest.mark.gpu
def test_node_run_loop_to_t_end_matches_python(tmp_path):
    """The reference's API run loop (src/ui/api.ts:669-699: encodeSteps -> submitted -> clock read,
    stepsResume when a batch took nothing) up to tEnd through the facade ends on the state and
    step count the ctypes caller's runToEnd reaches, bit for bit."""
    from shaderous_amd.sim import Sim, config_for
    pb = problems.kelvin_helmholtz(64)
    u1, u2 = pb.state(np.float32)
    job = _kh_job(pb, loop=True)
    job["cfg"]["tEnd"] = 0.05
    o1, o2, out = _run(tmp_path, job, u1, u2)
    sim = Sim(config_for(pb, t_end=0.05))
    try:
        sim.set_state(u1, u2)
        steps = sim.run_to_end()
        p1, p2 = sim.read_state()
        ck = sim.clock()
    finally:
        sim.destroy()
    assert out["clock"][3] == steps and out["clock"][0] == 0 and abs(out["clock"][1] - 0.05) < 1e-6
    assert sum(out["batches"]) >= steps
    assert np.array_equal(o1, p1) and np.array_equal(o2, p2)
    assert out["d1"]["t"] == ck[1] + ck[0]


@pytest.mark.gpu
def test_node_reset_reloads_the_initial_state(tmp_path):
    pb = problems.kelvin_helmholtz(64)
    u1, u2 = pb.state(np.float32)
    d = tmp_path / "r"
    _, _, out = _run(tmp_path, _kh_job(pb, reset=True), u1, u2)
    r1 = np.fromfile(tmp_path / "job" / "r1.bin", np.float32).reshape(-1, 4)
    assert np.array_equal(r1, u1)
    assert out["resetClock"][0] == 0 and out["resetClock"][1] == 0 and out["resetClock"][3] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "amr"])
def test_node_readbacks_pending_across_steps_regrid_and_destroy(tmp_path, kind):
    """Readbacks asked for and left pending while the sim steps, re-meshes and is destroyed each
    resolve with the state as it was when asked for (the copy is enqueued at the call, as
    mapAsync's is); a second readDiagnostics() while one is in flight resolves null."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import Sim, SimConfig, config_for
    if kind == "uniform":
        pb = problems.kelvin_helmholtz(64)
        u1, u2 = pb.state(np.float32)
        job = _kh_job(pb, pending=True)
        sim = Sim(config_for(pb))
        leaf = slice(None)
    else:
        blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))
        new = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c in ((0, 0, 0), (1, 0, 0)))

        def ic(x, y, z):
            return {"rho": 1.0 + 0.5 * np.exp(-(x * x + y * y + z * z) / 0.02), "vx": 0.2, "p": 1.0}
        u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, ic, 1.4)
        job = {"kind": "amr", "steps": 2, "after": 2, "regrid": new.reshape(-1).tolist(), "pending": True,
               "cfg": {"n": 32, "ndim": 3, "limiter": "mc", "isentropicRecon": False, "cfl": 0.4, "box": 1.0},
               "tree": {"root": [2, 2, 2], "blocks": blocks.reshape(-1).tolist(), "path": "fused"}}
        cfg = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4)
        sim = AmrSim(cfg, blocks, (2, 2, 2))
        leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    d = tmp_path / "job"
    d.mkdir()
    import json as _json
    (d / "job.json").write_text(_json.dumps(job))
    u1.tofile(d / "u1.bin")
    u2.tofile(d / "u2.bin")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "steps.js"), str(d)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    out = _json.loads((d / "out.json").read_text())
    assert out["secondWasNull"] is True
    b1 = np.fromfile(d / "b1.bin", np.float32).reshape(-1, 4)
    a1 = np.fromfile(d / "a1.bin", np.float32).reshape(-1, 4)
    o1 = np.fromfile(d / "o1.bin", np.float32).reshape(-1, 4)
    assert np.array_equal(b1[leaf], u1[leaf])
    try:
        sim.set_state(u1, u2)
        sim.step(job["steps"])
        p1, _ = sim.read_state()
        assert np.array_equal(a1[leaf], p1[leaf])
        if kind == "amr":
            sim.regrid(new)
            sim.step(job["after"])
            q1, _ = sim.read_state()
            nl = amr_tree.leaf_mask(new).repeat(4096)
            assert o1.shape == q1.shape and np.array_equal(o1[nl], q1[nl])
        else:
            assert np.array_equal(o1, p1)
        assert abs(out["d0"]["mass"] - sim.diagnostics().mass) < 1e-6 * out["d0"]["mass"]
    finally:
        sim.destroy()
