'use strict';
/**
 * The reference's simulation API for the hydro path (src/sim.ts `Simulation` :465-627, `Sim`,
 * src/amr/amr_sim.ts `AmrSim`, src/create.ts `createSim`), backed by libshx on an MI355X through
 * the N-API addon shx.node. Names and meanings follow the reference: SimConfig fields
 * (src/sim.ts:48-264) and configFor's defaults (:645-736); the stepping members the run loops
 * call (encodeSteps :485, submitted :492, stepsResume :503, src/ui/api.ts:669-699); step
 * (:3592), runToEnd (:3689-3718), readState / setState (:3790-3807, :3031-3065), diagnostics
 * and readDiagnostics (:1647-1673, :3598-3640), fluxMismatch (:3724-3787), profile
 * (:3541-3590), reset (:2704), writeParams (:2680), buildHydro (:2434), destroy (:3852-3876).
 * Members the reference makes asynchronous return Promises here too. Errors are thrown as
 * Error, as the reference does. There is no CPU path: without the addon or a GPU the calls throw.
 *
 * The GPU work of a step is enqueued on the sim's own HIP stream when encodeSteps() is called;
 * the command encoder the reference records into is accepted and not used.
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

/** domainLo (src/domain.ts:17-26): -box/2 on every axis of a Cartesian box. */
function domainLo(cfg) {
  return cfg.org ? cfg.org.slice(0, 3) : [-cfg.box / 2, -cfg.box / 2, -cfg.box / 2];
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
    org: domainLo(cfg),
  };
}

function live(h) {
  if (!h) throw new Error('the simulation was destroyed');
  return h;
}

/**
 * diagnosticsFrom (src/sim.ts:1647-1673): the Diagnostics record from the 24 reduction channels
 * (hydro/diagnostics.wgsl, sums in FP64) and the clock read with them.
 */
function diagnosticsFrom(r, clock) {
  const ranges = [];
  for (let i = 0; i < 5; i++) ranges.push([r[10 + 2 * i], r[11 + 2 * i]]);
  return {
    t: clock[1] + clock[0],
    dt: clock[0],
    steps: clock[3] + (clock[0] > 0 ? 1 : 0),
    held: clock[10] !== 0,
    mass: r[0],
    momentum: [r[1], r[2], r[3]],
    eKin: r[4],
    eInt: r[5],
    eGrav: r[6],
    eTot: r[4] + r[5] + r[6],
    maxMach: r[7],
    floored: r[8],
    jeansUnresolved: r[9],
    nuclearHeating: r[23],
    com: [r[20] / r[0], r[21] / r[0], r[22] / r[0]],
    perturber: null,
    masses: [],
    ranges,
  };
}

/**
 * What both grids share: the stepping members of `Simulation` and the readbacks. `n` is
 * the sim's native entry table (uniform or block).
 */
class Base {
  constructor(cfg, n) {
    this.cfg = cfg;
    this.n = n;
    this.h = null;
    this.initial = null;       // the state reset() loads: cfg.initialState, else the first setState()
    this.diagBusy = false;
    this.preparing = false;
  }

  /**
   * Encode up to `count` steps (Simulation.encodeSteps, src/sim.ts:485); returns how many were
   * enqueued. `enc` (the reference's GPUCommandEncoder) is not used: the steps go onto the
   * sim's stream. `mayHold` and `speculate` concern an adaptive mesh's own refinement checks,
   * which this path leaves to its caller: every step asked for is taken.
   */
  encodeSteps(enc, count, mayHold = false, speculate = 0) {
    return this.n.step(live(this.h), count);
  }

  /** Call after submitting an encoder encodeSteps() wrote to: the steps are already queued. */
  submitted() {}

  /** No step waits on a refinement check here: never a promise. */
  stepsResume() {
    return null;
  }

  step(count = 1) {
    this.encodeSteps(null, count);
  }

  /** Wait until every enqueued step is done. */
  sync() {
    this.n.sync(live(this.h));
  }

  clock() {
    return this.n.readClock(live(this.h));
  }

  /**
   * Replace the state with raw conserved arrays (Sim.setState, src/sim.ts:3031): the clock starts
   * at t = 0 (clockState). `rhoMean` serves self-gravity, which this path has none of.
   */
  setState(u1, u2, rhoMean = 0) {
    this.load(u1, u2, 0, 0);
  }

  /** The state and the clock it starts from (a restart: clockState(cfg, time, steps)). */
  load(u1, u2, time = 0, steps = 0) {
    this.n.setState(live(this.h), u1, u2, time, steps);
    if (!this.initial) this.initial = { u1: u1.slice(), u2: u2.slice() };
  }

  /**
   * Load the initial conditions (Simulation.reset): cfg.initialState ({u1, u2}) or the state the
   * first setState() loaded, with the clock of cfg.restart ({time, steps}) if given.
   */
  reset() {
    const ic = this.cfg.initialState || this.initial;
    if (!ic) throw new Error('reset() needs cfg.initialState or a state loaded by setState()');
    const r = this.cfg.restart || {};
    this.n.setState(live(this.h), ic.u1, ic.u2, r.time || 0, r.steps || 0);
  }

  /** Current conserved state (Sim.readState): the copies are enqueued now, after the steps so far. */
  async readState() {
    const r = await this.n.readStateAsync(live(this.h));
    return { u1: r.u1, u2: r.u2 };
  }

  readStateAsync() {
    return this.readState();
  }

  /** Reduce and read back the diagnostics (Sim.diagnostics). */
  async diagnostics() {
    const r = await this.n.diagnosticsAsync(live(this.h));
    return diagnosticsFrom(r.diag, r.clock);
  }

  /**
   * Reduce and read back the diagnostics; null if a readback is already in flight, so a frame
   * loop can call this freely without ever waiting on the GPU (Sim.readDiagnostics).
   */
  async readDiagnostics() {
    if (this.diagBusy) return null;
    this.diagBusy = true;
    try {
      return await this.diagnostics();
    } finally {
      this.diagBusy = false;
    }
  }

  destroy() {
    if (this.h) native.destroy(this.h);
    this.h = null;
  }
}

/** The uniform-grid simulation (the reference's `Sim`). */
class Sim extends Base {
  constructor(cfg) {
    super(cfg, {
      step: native.step,
      sync: native.sync,
      readClock: native.readClock,
      setState: native.setState,
      readStateAsync: native.readStateAsync,
      diagnosticsAsync: native.diagnosticsAsync,
    });
    this.h = native.create(toNative(cfg));
    const n = dims(cfg);
    this.nz = n[2];
    this.ncell = n[0] * n[1] * n[2];
    if (cfg.initialState) this.reset();
  }

  static create(cfg) {
    return new Sim(cfg);
  }

  /** Sim.runToEnd (src/sim.ts:3689-3718): the clock's step count once it holds at tEnd. */
  async runToEnd(batch = 16, maxSteps = 1e6) {
    if (!Number.isFinite(this.cfg.tEnd)) throw new Error('runToEnd needs a finite tEnd');
    return native.runToEnd(live(this.h), batch, Math.min(maxSteps, 0x7fffffff));
  }

  /**
   * Sim.writeParams (src/sim.ts:2680): the run-time fields of this.cfg (gamma, cfl, tEnd,
   * floors, etaDual, shockThresh, fluff) into the parameter block, from the next step on.
   */
  writeParams() {
    native.writeParams(live(this.h), toNative(this.cfg));
  }

  /** Sim.buildHydro (src/sim.ts:2434): the kernels of this.cfg's limiter, solver and reconstruction. */
  async buildHydro() {
    native.buildHydro(live(this.h), toNative(this.cfg));
  }

  /** Sim.fluxMismatch (src/sim.ts:3724-3787): faces whose two cells' mass fluxes differ. */
  async fluxMismatch() {
    const v = native.fluxMismatch(live(this.h));
    return { mismatched: v[0], faces: v[1] };
  }

  /** Sim.profile (src/sim.ts:3541-3590): milliseconds per pass. */
  async profile(steps = 10) {
    const ms = native.profile(live(this.h), steps);
    return { rates: ms[0], clock: ms[1], predict: ms[2], update: ms[3], step: ms[4] };
  }
}

/**
 * The block-AMR simulation (the reference's `AmrSim`) on a 2:1-balanced tree: `blocks` is an
 * Int32Array of (level, ix, iy, iz) rows, parents and leaves; `root` the root blocks per axis.
 */
class AmrSim extends Base {
  constructor(cfg, { root, blocks, path = 'fused', compensated = true }) {
    super(cfg, {
      step: native.amrStep,
      sync: native.amrSync,
      readClock: native.amrReadClock,
      setState: native.amrSetState,
      readStateAsync: native.readStateAsync,
      diagnosticsAsync: native.diagnosticsAsync,
    });
    const n = root.map((r) => 16 * r);
    const ncfg = toNative({ ...cfg, n, ndim: 3 });
    ncfg.dx = cfg.dx !== undefined ? cfg.dx : cfg.box / n[0];
    this.h = native.amrCreate(ncfg, { root, path: code(AMR_PATHS, path, 'AMR path'), compensated }, blocks);
    this.nz = n[2];
    this.nblocks = blocks.length / 4;
    this.ncell = this.nblocks * 4096;
    if (cfg.initialState) this.reset();
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

  /**
   * Re-mesh to `blocks` (Int32Array rows), within one level of the current tree everywhere. The
   * state reset() loads was of the old tree: it is dropped.
   */
  regrid(blocks) {
    native.amrRegrid(live(this.h), blocks);
    this.nblocks = blocks.length / 4;
    this.ncell = this.nblocks * 4096;
    this.initial = null;
  }

  async profile(steps = 5) {
    const ms = native.amrProfile(live(this.h), steps);
    return { assemble: ms[0], clock: ms[1], predict: ms[2], update: ms[3], step: ms[4] };
  }
}

/** createSim (src/create.ts:58-61): a block grid when a tree is given, else the uniform grid. */
function createSim(cfg, tree) {
  return tree ? AmrSim.create(cfg, tree) : Sim.create(cfg);
}

module.exports = { configFor, createSim, Sim, AmrSim, diagnosticsFrom, version: native.version, native };
