'use strict';
// Drives the hydro path through the Node binding, as a TypeScript caller of the reference's API
// would: node steps.js <dir>. Reads <dir>/job.json ({kind: 'uniform'|'amr', cfg, steps, tree?,
// regrid?, after?, loop?, pending?}) and the initial state <dir>/u1.bin, <dir>/u2.bin
// (little-endian floats of the run's precision), steps, and writes <dir>/o1.bin, <dir>/o2.bin,
// <dir>/out.json (clock, diagnostics).
//
// loop: the reference's API run loop (src/ui/api.ts:669-699) up to the config's tEnd: batches
// through encodeSteps(enc, n, mayHold) -> submitted() -> the clock read back without blocking
// (readClock, as the loop's clockCopy), stepsResume() when a batch took nothing, until the clock
// holds (dt = 0).
// probe: the probe() of that fractional position after the steps. toEnd: runToEnd() after them.
// tree.timeBins: per-block timesteps on the block grid.
// pending: readbacks left pending while the sim steps, re-meshes and is destroyed; each must
// still resolve with the state as it was when it was asked for.
const fs = require('fs');
const path = require('path');
const shx = require(path.join(__dirname, '..', '..', 'shaderous_amd', 'node'));

async function runLoop(sim, out) {
  let grow = 1;
  const batches = [];
  for (let guard = 0; guard < 100000; guard++) {
    const c = await sim.readClock();
    if (c[0] === 0 && c[1] > 0) break;
    const n = grow;
    const taken = sim.encodeSteps({}, n, false);
    sim.submitted();
    if (!taken) {
      const resume = sim.stepsResume();
      if (resume) await resume;
      continue;
    }
    batches.push(taken);
    grow = Math.min(grow * 2, 64);
  }
  out.batches = batches;
}

async function pending(sim, job, out) {
  // a state read, then steps and diagnostics asked for before it resolves
  const before = sim.readState();
  const diagBefore = sim.readDiagnostics();
  const second = await sim.readDiagnostics();     // null: a readback is in flight
  sim.step(job.steps);
  const after = sim.readState();
  if (job.regrid) {
    sim.regrid(Int32Array.from(job.regrid));
    sim.step(job.after || 0);
  }
  const last = sim.readState();
  const lastDiag = sim.diagnostics();
  sim.destroy();
  const [b, d0, a, l, ld] = await Promise.all([before, diagBefore, after, last, lastDiag]);
  out.secondWasNull = second === null;
  out.d0 = d0;
  out.d1 = ld;
  fs.writeFileSync(path.join(out.dir, 'b1.bin'), Buffer.from(b.u1.buffer));
  fs.writeFileSync(path.join(out.dir, 'a1.bin'), Buffer.from(a.u1.buffer));
  fs.writeFileSync(path.join(out.dir, 'o1.bin'), Buffer.from(l.u1.buffer));
  fs.writeFileSync(path.join(out.dir, 'o2.bin'), Buffer.from(l.u2.buffer));
}

async function main() {
  const dir = process.argv[2];
  const job = JSON.parse(fs.readFileSync(path.join(dir, 'job.json'), 'utf8'));
  const cfg = shx.configFor(job.cfg.n, job.cfg);
  const Arr = cfg.precision === 64 ? Float64Array : Float32Array;
  const load = (f) => {
    const b = fs.readFileSync(path.join(dir, f));
    return new Arr(b.buffer, b.byteOffset, b.length / Arr.BYTES_PER_ELEMENT).slice();
  };
  let sim;
  if (job.kind === 'amr') {
    const blocks = Int32Array.from(job.tree.blocks);
    sim = shx.createSim(cfg, { ...job.tree, blocks, path: job.tree.path || 'passes' });
  } else {
    sim = shx.createSim(cfg);
  }
  const out = { dir, version: shx.version() };
  sim.setState(load('u1.bin'), load('u2.bin'));
  if (job.pending) {
    await pending(sim, job, out);
    delete out.dir;
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify(out));
    console.log('node pending ok');
    return;
  }
  try {
    out.d0 = await sim.diagnostics();
    if (job.loop) {
      await runLoop(sim, out);
    } else {
      const taken = sim.encodeSteps({}, job.steps);
      sim.submitted();
      if (taken !== job.steps) throw new Error(`encodeSteps took ${taken} of ${job.steps}`);
      if (sim.stepsResume() !== null) throw new Error('stepsResume() is not null');
    }
    if (job.regrid) {
      // re-mesh, then step on the new tree
      sim.regrid(Int32Array.from(job.regrid));
      sim.step(job.after || 0);
    }
    const st = await sim.readState();
    const d1 = await sim.readDiagnostics();
    fs.writeFileSync(path.join(dir, 'o1.bin'), Buffer.from(st.u1.buffer));
    fs.writeFileSync(path.join(dir, 'o2.bin'), Buffer.from(st.u2.buffer));
    out.clock = Array.from(sim.clock());
    if (job.kind === 'amr') out.blocks = Array.from(sim.blocks);
    out.d1 = d1;
    if (job.probe) out.probe = await sim.probe(job.probe);
    if (job.kind !== 'amr') out.flux = await sim.fluxMismatch();
    if (job.toEnd) {
      out.toEnd = await sim.runToEnd();
      const e = await sim.readState();
      fs.writeFileSync(path.join(dir, 'e1.bin'), Buffer.from(e.u1.buffer));
    }
    if (job.reset) {
      // reset() reloads the first state and its clock
      sim.reset();
      const r = await sim.readState();
      fs.writeFileSync(path.join(dir, 'r1.bin'), Buffer.from(r.u1.buffer));
      out.resetClock = Array.from(sim.clock());
    }
    delete out.dir;
    fs.writeFileSync(path.join(dir, 'out.json'), JSON.stringify(out));
  } finally {
    sim.destroy();
  }
  let threw = false;
  try {
    sim.step(1);
  } catch (e) {
    threw = /destroyed/.test(e.message);
  }
  if (!threw) throw new Error('a destroyed simulation did not throw');
  console.log('node steps ok');
}

main().catch((e) => {
  console.error(e.stack || String(e));
  process.exit(1);
});
