'use strict';
// Drives the hydro path through the Node binding, as a TypeScript caller of the reference's API
// would: node steps.js <dir>. Reads <dir>/job.json ({kind: 'uniform'|'amr', cfg, steps, tree?,
// regrid?, after?})
// and the initial state <dir>/u1.bin, <dir>/u2.bin (little-endian floats of the run's
// precision), steps, and writes <dir>/o1.bin, <dir>/o2.bin, <dir>/out.json (clock, diagnostics).
const fs = require('fs');
const path = require('path');
const shx = require(path.join(__dirname, '..', '..', 'shaderous_amd', 'node'));

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
    sim = shx.createSim(cfg, { root: job.tree.root, blocks, path: job.tree.path || 'passes' });
  } else {
    sim = shx.createSim(cfg);
  }
  try {
    sim.setState(load('u1.bin'), load('u2.bin'));
    const d0 = sim.diagnostics();
    sim.step(job.steps);
    if (job.regrid) {
      // re-mesh, then step on the new tree
      sim.regrid(Int32Array.from(job.regrid));
      sim.step(job.after || 0);
    }
    const st = await sim.readStateAsync();
    const d1 = await (sim.readDiagnostics ? sim.readDiagnostics() : Promise.resolve(sim.diagnostics()));
    fs.writeFileSync(path.join(dir, 'o1.bin'), Buffer.from(st.u1.buffer));
    fs.writeFileSync(path.join(dir, 'o2.bin'), Buffer.from(st.u2.buffer));
    const out = { clock: Array.from(sim.clock()), d0, d1, version: shx.version() };
    if (job.kind !== 'amr') out.flux = sim.fluxMismatch();
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
