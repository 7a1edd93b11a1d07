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
           "readClock", "diagnostics", "diagnosticsAsync", "clockAsync", "fluxMismatch", "profile", "amrCreate", "amrTable",
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


# Every method of the reference's Simulation (src/sim.ts:465-627, optional members aside), and its
# fields and getters
SIM_METHODS = ["encodeSteps", "submitted", "stepsResume", "readDiagnostics", "diagnostics", "runToEnd", "reset",
               "writeParams", "buildHydro", "refreshPotential", "encodeBrush", "placeStar", "settleStar",
               "massChanged", "addMass", "readMasses", "removeAddedMasses", "setTracerFixed", "removeMasses",
               "readGasLumps", "retuneMasses", "probe", "setSources", "destroy", "fieldResources", "status",
               "debris", "debrisLedger"]
SIM_GETTERS = ["device", "paramBuffer", "clockBuffer", "massCount", "displayDims", "fieldShader", "speciesStride"]
SIM_ASYNC = ["readState", "readStateAsync", "diagnostics", "readDiagnostics", "runToEnd", "buildHydro", "placeStar",
             "settleStar", "readMasses", "readGasLumps", "probe", "debris", "debrisLedger", "readClock"]


def test_facade_serves_the_simulation_members():
    """Both facades carry every member of the reference's Simulation (src/sim.ts:465-627), so a
    caller assigns them to it with no cast; the ones the reference makes asynchronous return
    promises; the members of subsystems this path has none of answer as the reference's Sim does
    without them, and those needing WebGPU objects or other physics throw."""
    out = json.loads(node(
        "const s=require('./shaderous_amd/node');"
        f"const M={json.dumps(SIM_METHODS)}, G={json.dumps(SIM_GETTERS)}, A={json.dumps(SIM_ASYNC)};"
        "const has=(C,n)=>{for(let p=C.prototype;p;p=Object.getPrototypeOf(p)){"
        "const d=Object.getOwnPropertyDescriptor(p,n);if(d)return d;}return null;};"
        "const miss=[];for(const C of [s.Sim,s.AmrSim]){for(const n of M){const d=has(C,n);"
        "if(!d||typeof d.value!=='function')miss.push(C.name+'.'+n);}"
        "for(const n of G){const d=has(C,n);if(!d||!d.get)miss.push(C.name+'.'+n);}"
        "for(const n of A){const d=has(C,n);if(!d||d.value.constructor.name!=='AsyncFunction')miss.push(C.name+'.'+n+' async');}}"
        "const only=['fluxMismatch','profile'].filter(n=>typeof s.Sim.prototype[n]!=='function');"
        "const b=Object.create(s.Sim.prototype);"
        "const thrown=[];for(const n of ['device','paramBuffer','clockBuffer','fieldShader'])"
        "{try{b[n];}catch(e){thrown.push(n);}}"
        "for(const f of [()=>b.encodeBrush({},{}),()=>b.fieldResources(0,0),()=>b.setSources([{}])])"
        "{try{f();}catch(e){thrown.push(/outside the hydro path/.test(e.message));}}"
        "Promise.all([b.readMasses(),b.readGasLumps(),b.debris(),b.debrisLedger(),b.placeStar({}).catch(e=>'threw')])"
        ".then(v=>console.log(JSON.stringify({miss,only,thrown,v,mass:b.massCount,add:b.addMass({}),"
        "stride:b.speciesStride,empty:b.setSources([])===undefined})));"))
    assert out["miss"] == [] and out["only"] == [], out
    assert out["thrown"] == ["device", "paramBuffer", "clockBuffer", "fieldShader", True, True, True], out
    assert out["v"] == [None, None, None, None, "threw"]
    assert out["mass"] == 0 and out["add"] is False and out["stride"] == 0 and out["empty"] is True


def test_facade_takes_the_reference_config_and_refuses_other_physics():
    """A configuration shaped as the reference's SimConfig (the box as problem.box, fluff as
    { rho, u, v }, numeric boundary codes, the fields of physics this path does not run left
    off) maps to the native config this module's configFor() gives; each physics option outside
    the path is refused before anything touches a GPU, naming it."""
    out = json.loads(node(
        "const s=require('./shaderous_amd/node');"
        "const ours=s.configFor(32,{ndim:2,box:2,bcLo:['periodic','outflow',0],bcHi:['periodic','reflect',0],"
        "fluff:[0.01,2,3],limiter:'mc',tEnd:0.5});"
        "const ref={n:32,ndim:2,gext:[0,0,0],problem:{box:2},eos:'ideal',eosParams:{gamma:1.4,K:0,R:1},"
        "riemann:'hllc',limiter:'mc',gravity:'off',G:0,bcLo:[0,1,0],bcHi:[0,2,0],cfl:0.8,tEnd:0.5,"
        "rhoFloor:1e-10,pFloor:1e-14,etaDual:1e-3,shockThresh:0.5,fluff:{rho:0.01,u:2,v:3},"
        "isentropicRecon:true,cooling:'off',viscosity:0,conductivity:0,resistivity:0,radiation:'off',"
        "pressure:'explicit',reference:'off',cutCells:false,timeBins:0,driving:{accel:0},dampRate:0,"
        "startDamping:{rate:0,time:0},symmetry:[false,false,false]};"
        "const same=JSON.stringify(s.toNative(ours))===JSON.stringify(s.toNative(ref));"
        "const bad=[{gravity:'constant'},{gext:[0,-1,0]},{eos:'barotropic'},{cooling:'tabulated'},{viscosity:1e-3},"
        "{radiation:'m1'},{pressure:'implicit'},{bcLo:[4,1,0]},{riemann:'hllclm'},"
        "{problem:{box:2,geometry:'spherical'}},{fluff:{rho:1,u:1,v:1,va:2}}];"
        "const why=bad.map(o=>{try{new s.Sim({...ref,...o});return 'created';}catch(e){return e.message;}});"
        "console.log(JSON.stringify({same,why,a:s.toNative(ref)}))"))
    assert out["same"], out["a"]
    assert out["a"]["dx"] == 2 / 32 and out["a"]["org"] == [-1, -1, -1] and out["a"]["fluff"] == [0.01, 2, 3]
    for m in out["why"]:
        assert m != "created" and ("does not run this configuration" in m or "unknown riemann" in m), m


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


@pytest.mark.gpu
def test_node_probe_reads_the_cell(tmp_path):
    """probe(f) (src/sim.ts:3455-3480) after three steps: the cell at fraction f of the box, its
    centre relative to the box centre and its state, which equal the ctypes caller's cell."""
    from shaderous_amd.sim import Sim, config_for
    pb = problems.kelvin_helmholtz(64)
    u1, u2 = pb.state(np.float32)
    f = [0.3, 0.71, 0.5]
    _, _, out = _run(tmp_path, _kh_job(pb, probe=f), u1, u2)
    with Sim(config_for(pb)) as sim:
        sim.set_state(u1, u2)
        sim.step(3)
        p1, p2 = sim.read_state()
    i, j = int(f[0] * 64), int(f[1] * 64)
    pr = out["probe"]
    assert pr["index"] == [i, j]
    assert np.allclose(pr["pos"], [(i + 0.5) * pb.dx - pb.box / 2, (j + 0.5) * pb.dx - pb.box / 2, 0.0])
    c = i + 64 * j
    assert pr["u1"] == p1[c].tolist() and pr["u2"] == p2[c].tolist()


@pytest.mark.gpu
def test_node_amr_run_to_end_and_probe_match_python(tmp_path):
    """AmrSim.runToEnd (src/amr/amr_sim.ts:3749-3775) through the facade returns the ctypes
    caller's step count and state; its probe finds the leaf cell at a position on the refined
    block (its index at that block's level)."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    blocks = amr_tree.build_tree((2, 2, 2), 1, lambda L, c: c == (0, 0, 0))

    def ic(x, y, z):
        return {"rho": 1.0 + 0.5 * np.exp(-(x * x + y * y + z * z) / 0.02), "vx": 0.2, "p": 1.0}
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, ic, 1.4)
    f = [0.1, 0.2, 0.3]
    job = {"kind": "amr", "steps": 2, "probe": f, "toEnd": True,
           "cfg": {"n": 32, "ndim": 3, "limiter": "mc", "isentropicRecon": False, "cfl": 0.4, "box": 1.0,
                   "tEnd": 0.01},
           "tree": {"root": [2, 2, 2], "blocks": blocks.reshape(-1).tolist(), "path": "fused"}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4, t_end=0.01)
    sim = AmrSim(cfg, blocks, (2, 2, 2))
    try:
        sim.set_state(u1, u2)
        sim.step(2)
        p1, p2 = sim.read_state()
        steps = sim.run_to_end()
        e1, _ = sim.read_state()
    finally:
        sim.destroy()
    assert out["toEnd"] == steps
    leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    n1 = np.fromfile(tmp_path / "job" / "e1.bin", np.float32).reshape(-1, 4)
    assert np.array_equal(n1[leaf], e1[leaf])
    pr = out["probe"]
    idx = [int(x * 64) for x in f]           # the refined root block's children cover [0, 0.5) of the box
    assert pr["index"] == idx and pr["block"]["level"] == 1 and pr["block"]["coords"] == [x // 16 for x in idx]
    assert tuple(blocks[pr["block"]["slot"]]) == (1, *[x // 16 for x in idx])
    slot = pr["block"]["slot"]
    loc = [x % 16 for x in idx]
    cell = 4096 * slot + loc[0] + 16 * (loc[1] + 16 * loc[2])
    assert pr["u1"] == p1[cell].tolist() and pr["u2"] == p2[cell].tolist()
    assert np.allclose(pr["pos"], [(x + 0.5) / 64 - 0.5 for x in idx])


@pytest.mark.gpu
def test_node_amr_time_bins_match_python(tmp_path):
    """Per-block timesteps through the facade (tree.timeBins, the reference's cfg.timeBins): the
    Node caller's cycles give the ctypes caller's bits."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    from test_gpu_bins import flow
    from test_oracle_bins import refined_tree
    blocks = refined_tree()
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 32, flow, 1.4)
    job = {"kind": "amr", "steps": 8, "cfg": {"n": 32, "ndim": 3, "limiter": "mc", "isentropicRecon": False,
                                              "cfl": 0.4, "box": 1.0},
           "tree": {"root": [2, 2, 2], "blocks": blocks.reshape(-1).tolist(), "timeBins": 2}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(32, 32, 32), dx=1.0 / 32, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, (2, 2, 2), path="passes", time_bins=2)
    try:
        sim.set_state(u1, u2)
        sim.step(8)
        p1, p2 = sim.read_state()
        ck = sim.clock()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks).repeat(4096)
    assert np.array_equal(o1[leaf], p1[leaf]) and np.array_equal(o2[leaf], p2[leaf])
    assert out["clock"][3] == ck[3]


@pytest.mark.gpu
def test_node_amr_2d_matches_python(tmp_path):
    """A quadtree (cfg.ndim 2) through the facade: the Node caller's steps give the ctypes
    caller's bits, and its probe finds the leaf cell at a position in the refined patch."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    from test_gpu_amr2d import ROOT, flow, quadtree
    blocks = quadtree()
    u1, u2 = amr_tree.blocks_state(blocks, 1.0 / 64, flow, 1.4, ndim=2)
    job = {"kind": "amr", "steps": 4, "cfg": {"n": 64, "ndim": 2, "limiter": "mc", "isentropicRecon": False,
                                              "cfl": 0.4, "box": 1.0},
           "tree": {"root": list(ROOT), "blocks": blocks.reshape(-1).tolist()}, "probe": [0.44, 0.45, 0.5]}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(64, 64, 1), ndim=2, dx=1.0 / 64, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, ROOT, path="passes")
    try:
        sim.set_state(u1, u2)
        sim.step(4)
        p1, p2 = sim.read_state()
    finally:
        sim.destroy()
    leaf = amr_tree.leaf_mask(blocks).repeat(256)
    assert np.array_equal(o1[leaf], p1[leaf]) and np.array_equal(o2[leaf], p2[leaf])
    pr = out["probe"]
    assert pr["block"]["level"] == 2 and pr["index"][2] == 0 and pr["pos"][2] == 0
    L, bx, by, _ = blocks[pr["block"]["slot"]]
    idx = pr["index"]
    assert idx == [int(0.44 * (64 << 2)), int(0.45 * (64 << 2)), 0] and (bx, by) == (idx[0] // 16, idx[1] // 16)
    cell = pr["block"]["slot"] * 256 + (idx[0] - 16 * bx) + 16 * (idx[1] - 16 * by)
    assert np.array_equal(np.float32(pr["u1"]), p1[cell]) and np.array_equal(np.float32(pr["u2"]), p2[cell])


@pytest.mark.gpu
def test_node_amr_regrid_cadence_matches_python(tmp_path):
    """The re-mesh cadence through the facade (tree.regridEvery, on by default where the
    reference's config has amrLevels): the Node caller's 12 steps re-mesh as the ctypes caller's
    sim with regrid_every does, to the same tree and bits, and the facade follows the tree."""
    from shaderous_amd.amr import AmrSim
    from shaderous_amd.sim import SimConfig
    blocks, dx0 = amr_tree.blast_tree(root=4, max_level=2, radius=0.2, center=(0.03, -0.02, 0.05))

    def blast(x, y, z):
        r = np.sqrt((x - 0.03) ** 2 + (y + 0.02) ** 2 + (z - 0.05) ** 2)
        return {"rho": 1.0 + 0.0 * x, "p": np.where(r < 0.2, 100.0, 1.0)}
    u1, u2 = amr_tree.blocks_state(blocks, dx0, blast, 1.4)
    job = {"kind": "amr", "steps": 12, "cfg": {"n": 64, "ndim": 3, "limiter": "mc", "isentropicRecon": False,
                                               "cfl": 0.4, "box": 1.0, "amrLevels": 2, "eosParams": {"gamma": 1.4}},
           "tree": {"root": [4, 4, 4], "blocks": blocks.reshape(-1).tolist(), "path": "fused", "regridEvery": 4,
                    "refineAbove": 0.05, "refineBelow": 0.01}}
    o1, o2, out = _run(tmp_path, job, u1, u2)
    cfg = SimConfig(n=(64, 64, 64), dx=1.0 / 64, gamma=1.4, limiter="mc", isentropic=False, cfl=0.4)
    sim = AmrSim(cfg, blocks, (4, 4, 4), path="fused", regrid_every=4, max_level=2, refine_above=0.05,
                 coarsen_below=0.01)
    try:
        sim.set_state(u1, u2)
        sim.step(12)
        p1, p2 = sim.read_state()
        tree = sim.blocks.copy()
    finally:
        sim.destroy()
    assert len(tree) != len(blocks) or not np.array_equal(tree, blocks)
    assert np.array_equal(np.asarray(out["blocks"], np.int32).reshape(-1, 4), tree)
    leaf = amr_tree.leaf_mask(tree).repeat(4096)
    assert np.array_equal(o1[leaf], p1[leaf]) and np.array_equal(o2[leaf], p2[leaf])
