'use strict';
/**
 * The reference's simulation API for the hydro path (src/sim.ts `Simulation`/`Sim`, src/amr/amr_sim.ts
 * `AmrSim`, src/create.ts `createSim`), backed by libshx on an MI355X through the N-API addon
 * shx.node. Names and meanings follow the reference: SimConfig fields (src/sim.ts:48-264) and
 * configFor's defaults (:645-736), step (:3592), runToEnd (:3689-3718), readState / setState
 * (:3790-3807, :3031-3065), diagnostics (:1647-1673), fluxMismatch (:3724-3787), profile
 * (:3541-3590), destroy (:3852-3876). Errors are thrown as Error, as the reference does.
 * There is no CPU path: without the addon or a GPU the calls throw.
 */
const native = require('./shx.node');

const LIMITERS = { koren: 0, mc: 1, vanleer: 2, minmod: 3 };
const RIEMANN = { hllc: 0, hll: 1 };
const BC = { periodic: 0, outflow: 1, reflect: 2, reflecting: 2, 'outflow-only': 3 };
const PATHS = { fused: 0, plain: 1 };
const AMR_PATHS = { fused: 0, passes: 1, levels: 2 };

function code(table, v, what) {
  if (typeof v === 'number') return v;
  if (!(v in table)) throw new Error(`unknown ${what} '${v}'`);
  return table[v];
}

/**
 * configFor (src/sim.ts:645-736) for the hydro path: an n^ndim grid on a box of side `box`
 * (cells centred at (i + 1/2) dx - box/2, src/domain.ts:1-40), with overrides.
 */
function configFor(n, over = {}) {
  const cfg = {
    n,
    ndim: 3,
    box: 1,
    bcLo: [0, 0, 0],
    bcHi: [0, 0, 0],
    riemann: 'hllc',
    limiter: 'koren',
    isentropicRecon: true,
    eosParams: { gamma: 1.4 },
    cfl: 0.8,
    tEnd: Infinity,
    rhoFloor: 1e-10,
    pFloor: 1e-14,
    etaDual: 1e-3,
    shockThresh: 0.5,
    fluff: [0, 0, 0],
    precision: 32,
    path: 'fused',
    device: 0,
    tile: [0, 0, 0],
    ...over,
  };
  if (over.eosParams) cfg.eosParams = { gamma: 1.4, ...over.eosParams };
  return cfg;
}

function dims(cfg) {
  if (Array.isArray(cfg.n)) return cfg.n.slice(0, 3);
  if (cfg.ndim === 2) return [cfg.n, cfg.n, 1];
  if (cfg.ndim === 1) return [cfg.n, 1, 1];
  return [cfg.n, cfg.n, cfg.n];
}

function toNative(cfg) {
  const n = dims(cfg);
  return {
    n,
    ndim: cfg.ndim,
    bcLo: cfg.bcLo.map((b) => code(BC, b, 'boundary')),
    bcHi: cfg.bcHi.map((b) => code(BC, b, 'boundary')),
    limiter: code(LIMITERS, cfg.limiter, 'limiter'),
    riemann: code(RIEMANN, cfg.riemann, 'riemann solver'),
    isentropic: !!cfg.isentropicRecon,
    precision: cfg.precision,
    path: code(PATHS, cfg.path, 'path'),
    device: cfg.device,
    dx: cfg.dx !== undefined ? cfg.dx : cfg.box / n[0],
    gamma: cfg.eosParams.gamma,
    cfl: cfg.cfl,
    rhoFloor: cfg.rhoFloor,
    pFloor: cfg.pFloor,
    etaDual: cfg.etaDual,
    shockThresh: cfg.shockThresh,
    tEnd: Number.isFinite(cfg.tEnd) ? cfg.tEnd : 1e30,
    fluff: cfg.fluff,
    tile: cfg.tile,
  };
}

function live(h) {
  if (!h) throw new Error('the simulation was destroyed');
  return h;
}

/** diagnosticsFrom (src/sim.ts:1647-1673): the Diagnostics record, sums in FP64. */
function diagnosticsFrom(r, clock) {
  return {
    t: clock[1] + clock[0],
    dt: clock[0],
    steps: clock[3] + (clock[0] > 0 ? 1 : 0),
    held: clock[10] !== 0,
    mass: r[0],
    momentum: [r[1], r[2], r[3]],
    eKin: r[4],
    eInt: r[5],
    eGrav: 0,
    eTot: r[6],
    maxMach: r[7],
    floored: r[8],
    scalar: r[9],
    ranges: [[r[10], r[11]], [r[12], r[13]]],
  };
}

/** The uniform-grid simulation (the reference's `Sim`). */
class Sim {
  constructor(cfg) {
    this.cfg = cfg;
    this.h = native.create(toNative(cfg));
    const n = dims(cfg);
    this.ncell = n[0] * n[1] * n[2];
  }

  static create(cfg) {
    return new Sim(cfg);
  }

  setState(u1, u2, time = 0, steps = 0) {
    native.setState(live(this.h), u1, u2, time, steps);
  }

  readState() {
    return native.readState(live(this.h));
  }

  readStateAsync() {
    return native.readStateAsync(live(this.h));
  }

  step(count = 1) {
    native.step(live(this.h), count);
  }

  submitted() {
    native.sync(live(this.h));
  }

  runToEnd(batch = 64, maxSteps = 1 << 30) {
    return native.runToEnd(live(this.h), batch, maxSteps);
  }

  clock() {
    return native.readClock(live(this.h));
  }

  diagnostics() {
    return diagnosticsFrom(native.diagnostics(live(this.h)), native.readClock(live(this.h)));
  }

  async readDiagnostics() {
    const r = await native.diagnosticsAsync(live(this.h));
    return diagnosticsFrom(r, native.readClock(live(this.h)));
  }

  fluxMismatch() {
    const v = native.fluxMismatch(live(this.h));
    return { mismatched: v[0], faces: v[1] };
  }

  profile(steps = 10) {
    const ms = native.profile(live(this.h), steps);
    return { rates: ms[0], clock: ms[1], predict: ms[2], update: ms[3], step: ms[4] };
  }

  destroy() {
    if (this.h) native.destroy(this.h);
    this.h = null;
  }
}

/**
 * The block-AMR simulation (the reference's `AmrSim`) on a fixed 2:1-balanced tree: `blocks` is an
 * Int32Array of (level, ix, iy, iz) rows, parents and leaves; `root` the root blocks per axis.
 */
class AmrSim {
  constructor(cfg, { root, blocks, path = 'fused', compensated = true }) {
    this.cfg = cfg;
    const n = root.map((r) => 16 * r);
    const ncfg = toNative({ ...cfg, n, ndim: 3 });
    ncfg.dx = cfg.dx !== undefined ? cfg.dx : cfg.box / n[0];
    this.h = native.amrCreate(ncfg, { root, path: code(AMR_PATHS, path, 'AMR path'), compensated }, blocks);
    this.nblocks = blocks.length / 4;
    this.ncell = this.nblocks * 4096;
  }

  static create(cfg, tree) {
    return new AmrSim(cfg, tree);
  }

  table() {
    return native.amrTable(live(this.h));
  }

  stats() {
    const v = native.amrStats(live(this.h));
    return { blocks: v[0], leaves: v[1], matchLeaves: v[2], levels: v[3], launchesPerStep: v[4] };
  }

  /** Per block the largest relative density jump between face-adjacent cells (NaN for parents). */
  indicator() {
    return native.amrIndicator(live(this.h));
  }

  /** Re-mesh to `blocks` (Int32Array rows), within one level of the current tree everywhere. */
  regrid(blocks) {
    native.amrRegrid(live(this.h), blocks);
    this.nblocks = blocks.length / 4;
    this.ncell = this.nblocks * 4096;
  }

  setState(u1, u2, time = 0, steps = 0) {
    native.amrSetState(live(this.h), u1, u2, time, steps);
  }

  readState() {
    return native.amrReadState(live(this.h));
  }

  readStateAsync() {
    return native.readStateAsync(live(this.h));
  }

  step(count = 1) {
    native.amrStep(live(this.h), count);
  }

  submitted() {
    native.amrSync(live(this.h));
  }

  clock() {
    return native.amrReadClock(live(this.h));
  }

  diagnostics() {
    return diagnosticsFrom(native.amrDiagnostics(live(this.h)), native.amrReadClock(live(this.h)));
  }

  profile(steps = 5) {
    const ms = native.amrProfile(live(this.h), steps);
    return { assemble: ms[0], clock: ms[1], predict: ms[2], update: ms[3], step: ms[4] };
  }

  destroy() {
    if (this.h) native.destroy(this.h);
    this.h = null;
  }
}

/** createSim (src/create.ts:58-61): a block grid when a tree is given, else the uniform grid. */
function createSim(cfg, tree) {
  return tree ? AmrSim.create(cfg, tree) : Sim.create(cfg);
}

module.exports = { configFor, createSim, Sim, AmrSim, diagnosticsFrom, version: native.version, native };
