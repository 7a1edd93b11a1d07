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

/** The box side: this facade's `box`, or the reference's `problem.box` (src/problems/types.ts:258). */
function boxOf(cfg) {
  const box = cfg.box !== undefined ? cfg.box : cfg.problem && cfg.problem.box;
  if (!(box > 0) && cfg.dx === undefined) throw new Error('the configuration needs a box side (box or problem.box) or dx');
  return box;
}

/** domainLo (src/domain.ts:17-26): -box/2 on every axis of a Cartesian box. */
function domainLo(cfg) {
  if (cfg.org) return cfg.org.slice(0, 3);
  const box = cfg.box !== undefined || (cfg.problem && cfg.problem.box !== undefined) ? boxOf(cfg) : dims(cfg)[0] * cfg.dx;
  return [-box / 2, -box / 2, -box / 2];
}

/**
 * The tenuous-gas limits as (rho, u, v): this facade's triple, or the reference's
 * { rho, u, v } (SimConfig.fluff, src/sim.ts:187-191).
 */
function fluffOf(f) {
  if (!f) return [0, 0, 0];
  if (Array.isArray(f)) return f.slice(0, 3);
  return [f.rho || 0, f.u || 0, f.v || 0];
}

/**
 * Why this backend cannot run a configuration, or null: in the manner of refUnsupported
 * (src/sim.ts:912-940), the physics outside the hydro step it serves (the ideal gas, explicit
 * pressure, Cartesian coordinates, boundary codes 0-3 of src/problems/boundaries.ts:20-28, no
 * source terms). Fields this facade's own configFor() lacks are read only when present, so the
 * reference's SimConfig and this facade's both pass through.
 */
function hipUnsupported(cfg) {
  const p = cfg.problem || {};
  if (cfg.eos !== undefined && cfg.eos !== 'ideal') return `the '${cfg.eos}' equation of state (the ideal gas only)`;
  if (cfg.gravity !== undefined && cfg.gravity !== 'off') return `gravity '${cfg.gravity}'`;
  if (cfg.gext && cfg.gext.some((g) => g !== 0)) return 'a uniform external field (gext)';
  if (cfg.pressure === 'implicit') return 'the implicit pressure solver';
  if (cfg.reference !== undefined && cfg.reference !== 'off') return 'a state stored as a deviation from a reference state';
  if (cfg.radiation !== undefined && cfg.radiation !== 'off') return 'radiation';
  if (cfg.cooling !== undefined && cfg.cooling !== 'off') return `cooling '${cfg.cooling}'`;
  if (cfg.viscosity > 0 || cfg.conductivity > 0 || cfg.resistivity > 0) return 'viscosity, conduction or resistivity';
  if (cfg.cutCells) return 'cut cells';
  if ((cfg.driving && cfg.driving.accel > 0) || cfg.dampRate > 0 || (cfg.startDamping && cfg.startDamping.rate > 0))
    return 'driving or damping';
  if (cfg.symmetry && cfg.symmetry.some(Boolean)) return 'mirror symmetry';
  if (cfg.orbitOmega) return 'a rotating frame';
  if (cfg.fluff && !Array.isArray(cfg.fluff) && cfg.fluff.va) return 'an Alfven-speed limit on tenuous gas';
  if (p.geometry && p.geometry !== 'cartesian') return `${p.geometry} coordinates`;
  if (p.mhd || (p.species && p.species.length) || p.cosmology || p.shear || p.frame) return 'a magnetic field, species, comoving coordinates, a shearing box or a moving frame';
  if ((p.bodies && p.bodies.length) || p.obstacleDistances || (p.moving && p.moving.length) || p.damping) return 'obstacles or damping zones';
  if (typeof p.masses === 'function' && p.masses(cfg.params || {}, cfg.G || 0, boxOf(cfg) / dims(cfg)[0], cfg.ndim).length > 0) return 'point masses';
  for (const b of [...cfg.bcLo, ...cfg.bcHi])
    if (typeof b === 'number' && (b < 0 || b > 3)) return `boundary code ${b} (periodic, outflow, reflecting and outflow-only only)`;
  return null;
}

function toNative(cfg) {
  const why = hipUnsupported(cfg);
  if (why) throw new Error(`the MI355X backend does not run this configuration: ${why}`);
  const n = dims(cfg);
  return {
    n,
    ndim: cfg.ndim,
    bcLo: cfg.bcLo.map((b) => code(BC, b, 'boundary')),
    bcHi: cfg.bcHi.map((b) => code(BC, b, 'boundary')),
    limiter: code(LIMITERS, cfg.limiter, 'limiter'),
    riemann: code(RIEMANN, cfg.riemann, 'riemann solver'),
    isentropic: !!cfg.isentropicRecon,
    precision: cfg.precision !== undefined ? cfg.precision : 32,
    path: code(PATHS, cfg.path !== undefined ? cfg.path : 'fused', 'path'),
    device: cfg.device !== undefined ? cfg.device : 0,
    dx: cfg.dx !== undefined ? cfg.dx : boxOf(cfg) / n[0],
    gamma: cfg.eosParams.gamma,
    cfl: cfg.cfl,
    rhoFloor: cfg.rhoFloor,
    pFloor: cfg.pFloor,
    etaDual: cfg.etaDual,
    shockThresh: cfg.shockThresh,
    tEnd: Number.isFinite(cfg.tEnd) ? cfg.tEnd : 1e30,
    fluff: fluffOf(cfg.fluff),
    tile: cfg.tile !== undefined ? cfg.tile : [0, 0, 0],
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

/** Thrown by the members of `Simulation` that serve physics or WebGPU resources this backend has not. */
function outside(member, what) {
  return new Error(`${member}: ${what} is outside the hydro path the MI355X backend runs`);
}

/**
 * What both grids share: every member of the reference's `Simulation` (src/sim.ts:465-627).
 * The stepping members and readbacks run on the GPU; the members of subsystems this path has
 * none of behave as the reference's Sim does when the subsystem is absent (no point masses, no
 * self-gravity, no sources: src/sim.ts:3087-3090, 3383-3386, 3400-3414, 3447-3451, 3482-3491), and
 * those that would need WebGPU objects or physics outside the path throw. `nat` is the sim's
 * native entry table (uniform or block).
 */
class Base {
  constructor(cfg, nat) {
    this.cfg = cfg;
    this.nat = nat;
    this.h = null;
    this.cur = 0;              // the native sim keeps its own parity; readState() hands out the current state
    this.speciesCur = 0;
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
    return this.nat.step(live(this.h), count);
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
    this.nat.sync(live(this.h));
  }

  /** The clock (dt, t, tc, steps, ...) after every step enqueued so far; waits for them. */
  clock() {
    return this.nat.readClock(live(this.h));
  }

  /**
   * The clock as a readback: enqueued now, after the steps so far, and resolved when it lands,
   * without blocking (the run loop's clockCopy, src/ui/api.ts:503-534).
   */
  async readClock() {
    const r = await this.nat.clockAsync(live(this.h));
    return r.clock;
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
    this.nat.setState(live(this.h), u1, u2, time, steps);
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
    this.nat.setState(live(this.h), ic.u1, ic.u2, r.time || 0, r.steps || 0);
  }

  /** Current conserved state (Sim.readState): the copies are enqueued now, after the steps so far. */
  async readState() {
    const r = await this.nat.readStateAsync(live(this.h));
    return { u1: r.u1, u2: r.u2 };
  }

  async readStateAsync() {
    return this.readState();
  }

  /** Reduce and read back the diagnostics (Sim.diagnostics). */
  async diagnostics() {
    const r = await this.nat.diagnosticsAsync(live(this.h));
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

  // ---- members of subsystems outside this path (src/sim.ts:465-627) ----

  /** Point masses (src/sim.ts:3414-3416): a run here holds none. */
  get massCount() {
    return 0;
  }
  /** The species texture's stride (src/sim.ts:3512-3514): no species, 0. */
  get speciesStride() {
    return 0;
  }
  get device() {
    throw outside('device', 'a WebGPU device');
  }
  get paramBuffer() {
    throw outside('paramBuffer', 'a WebGPU parameter buffer');
  }
  /** The clock lives on the GPU as well; clock() reads it. */
  get clockBuffer() {
    throw outside('clockBuffer', 'a WebGPU clock buffer (read the clock with clock())');
  }
  get fieldShader() {
    throw outside('fieldShader', 'the WebGPU display pass');
  }
  fieldResources(parity, speciesParity, radParity) {
    throw outside('fieldResources', 'the WebGPU display pass');
  }
  /** No self-gravity: nothing to refresh (src/sim.ts:3087-3088 returns when no field is on). */
  refreshPotential() {}
  encodeBrush(enc, op) {
    throw outside('encodeBrush', 'the editing brush');
  }
  async placeStar(op) {
    throw outside('placeStar', 'placing stars');
  }
  async settleStar(star) {
    throw outside('settleStar', 'settling stars');
  }
  /** The gas mass feeds self-gravity's mean density (src/sim.ts:3487-3491): none here. */
  massChanged(mass) {}
  /** No room for point masses: false, as the reference's Sim says when it is full (src/sim.ts:3404-3405). */
  addMass(m) {
    return false;
  }
  /** "null if there are none" (src/sim.ts:539-542). */
  async readMasses() {
    return null;
  }
  /** None were added (src/sim.ts:3447-3448). */
  removeAddedMasses() {}
  /** An index among no masses (src/sim.ts:3400-3402). */
  setTracerFixed(index, fixed) {}
  /** Only added masses go (src/sim.ts:3419-3421): none here. */
  removeMasses(indices) {}
  /** Gravitationally bound lumps (gravity/lumps.wgsl): no gravity, no read. */
  async readGasLumps() {
    return null;
  }
  retuneMasses() {}
  /** Injection sources (src/sim.ts:3376-3380): an empty list is the run's own. */
  setSources(list) {
    if (list && list.length) throw outside('setSources', 'injection sources');
  }
  /** No debris is tracked (src/sim.ts:3007-3013 without a debris buffer). */
  async debris() {
    return null;
  }
  async debrisLedger() {
    return null;
  }
  status() {
    return '';
  }

  /**
   * Sim.probe (src/sim.ts:3455-3480): the cell at fractional position `f` of the box, its centre,
   * index, dx and state. The stencil (neighbours, fields) is not served.
   */
  async probe(f, opts) {
    if (opts && opts.stencil) throw outside('probe', 'the stencil probe');
    const dd = this.displayDims;
    const cell = (x, a) => Math.min(dd[a] - 1, Math.max(0, Math.floor(x * dd[a])));
    return this.probeCell([cell(f[0], 0), cell(f[1], 1), cell(f[2], 2)]);
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
      clockAsync: native.clockAsync,
    });
    this.h = native.create(toNative(cfg));
    const n = dims(cfg);
    this.dims = n;
    this.n = n[0];
    this.nz = n[2];
    this.ncell = n[0] * n[1] * n[2];
    if (cfg.initialState) this.reset();
  }

  /** The grid the display samples: the grid itself (src/sim.ts:3494-3496). */
  get displayDims() {
    return this.dims.slice();
  }

  /** One cell of the current state (the probe of src/sim.ts:3455-3480 without its stencil). */
  async probeCell(index) {
    const { u1, u2 } = await this.readState();
    const [nx, ny] = this.dims;
    const c = index[0] + nx * (index[1] + ny * index[2]);
    const nat = toNative(this.cfg);
    const pos = index.map((i, a) => (i + 0.5) * nat.dx + nat.org[a]);
    if (this.nz === 1) pos[2] = 0;
    return {
      pos,
      index: this.nz === 1 ? index.slice(0, 2) : index.slice(),
      dx: nat.dx,
      u1: Array.from(u1.subarray(4 * c, 4 * c + 4)),
      u2: Array.from(u2.subarray(4 * c, 4 * c + 4)),
      stencil: undefined,
    };
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
 * cfg.ndim 2: a quadtree of 16^2-cell blocks (root[2] = 1, every iz 0) on the passes paths.
 */
class AmrSim extends Base {
  constructor(cfg, { root, blocks, path, compensated = true, timeBins, regridEvery, maxLevel, refineAbove, refineBelow }) {
    super(cfg, {
      step: native.amrStep,
      sync: native.amrSync,
      readClock: native.amrReadClock,
      setState: native.amrSetState,
      readStateAsync: native.readStateAsync,
      diagnosticsAsync: native.diagnosticsAsync,
      clockAsync: native.clockAsync,
    });
    this.ndim = cfg.ndim === 2 ? 2 : 3;
    const n = level0(root, this.ndim);
    const ncfg = toNative({ ...cfg, n, ndim: this.ndim });
    // per-block timesteps (SimConfig.timeBins, at most 5): cycles of 2^K substeps, on the passes
    // path (the fused march takes one dt and steps 3D blocks)
    const bins = Math.max(0, Math.min(5, Math.floor(timeBins !== undefined ? timeBins : cfg.timeBins || 0)));
    const kind = path !== undefined ? path : bins > 0 || this.ndim === 2 ? 'passes' : 'fused';
    // the re-mesh cadence (REGRID_EVERY = 8, src/amr/amr_sim.ts:73): on by default where the
    // reference's config asks for levels of refinement (amrLevels), with its thresholds
    // (refineHi, refineLo) on the same indicator (amr/flag.wgsl's Loehner estimate)
    const levels = cfg.amrLevels || 0;
    const every = regridEvery !== undefined ? regridEvery : levels > 0 ? 8 : 0;
    const lay = {
      root,
      path: code(AMR_PATHS, kind, 'AMR path'),
      compensated,
      timeBins: bins,
      regridEvery: every,
      maxLevel: maxLevel !== undefined ? maxLevel : levels,
      refineAbove: refineAbove !== undefined ? refineAbove : cfg.refineHi !== undefined ? cfg.refineHi : 0.8,
      coarsenBelow: refineBelow !== undefined ? refineBelow : cfg.refineLo !== undefined ? cfg.refineLo : 0.2,
    };
    this.regridEvery = every;
    this.h = native.amrCreate(ncfg, lay, blocks);
    this.root = root.slice(0, 3);
    this.dx0 = ncfg.dx;
    this.tree(blocks);
    if (cfg.initialState) this.reset();
  }

  /** Steps as Simulation.encodeSteps; a sim that re-meshes by itself reads its new tree back. */
  encodeSteps(enc, count, mayHold = false, speculate = 0) {
    const taken = super.encodeSteps(enc, count, mayHold, speculate);
    if (this.regridEvery > 0) this.syncTree();
    return taken;
  }

  /** The current tree's rows from the table (after a re-mesh made by the steps). */
  syncTree() {
    const nb = native.amrStats(live(this.h))[0];
    this.nblocks = nb;
    const t = native.amrTable(live(this.h));
    const rows = new Int32Array(4 * nb);
    for (let i = 0; i < nb; i++) for (let k = 0; k < 4; k++) rows[4 * i + k] = t[32 * i + k];
    // the state reset() loads was of the old tree: dropped once the tree changed
    if (rows.length !== this.blocks.length || rows.some((v, i) => v !== this.blocks[i])) this.initial = null;
    this.tree(rows);
  }

  /** The tree's rows, its finest cells per side (`n`, `nz`: AmrSim's, src/amr/amr_sim.ts:184). */
  tree(blocks) {
    this.blocks = Int32Array.from(blocks);
    this.nblocks = blocks.length / 4;
    this.ncell = this.nblocks * 16 ** this.ndim;
    let top = 0;
    for (let i = 0; i < this.nblocks; i++) top = Math.max(top, blocks[4 * i]);
    this.levels = top + 1;
    this.n = (16 * this.root[0]) << top;
    this.nz = this.ndim === 2 ? 1 : (16 * this.root[2]) << top;
  }

  /** The uniform volume the display samples: the finest grid, at most 256 a side (src/amr/amr_sim.ts:3780-3785). */
  get displayDims() {
    const d = this.root.map((r) => Math.min((16 * r) << (this.levels - 1), 256));
    if (this.ndim === 2) d[2] = 1;
    return d;
  }

  /**
   * AmrSim.probe (src/amr/amr_sim.ts:3457-3495): the leaf cell at fractional position `f` of the
   * box (its index counted at the leaf's level), its centre, dx and state.
   */
  async probe(f, opts) {
    if (opts && opts.stencil) throw outside('probe', 'the stencil probe');
    const b = this.blocks;
    const parent = new Set();
    for (let i = 0; i < this.nblocks; i++)
      if (b[4 * i] > 0) parent.add([b[4 * i] - 1, b[4 * i + 1] >> 1, b[4 * i + 2] >> 1, b[4 * i + 3] >> 1].join());
    for (let i = 0; i < this.nblocks; i++) {
      const L = b[4 * i];
      const c = [b[4 * i + 1], b[4 * i + 2], b[4 * i + 3]];
      if (parent.has([L, ...c].join())) continue;
      const index = [0, 1, 2].map((a) => {
        if (a === 2 && this.ndim === 2) return 0;
        const across = (16 * this.root[a]) << L;
        return Math.min(across - 1, Math.max(0, Math.floor(f[a] * across)));
      });
      if (!index.every((x, a) => Math.floor(x / 16) === c[a])) continue;
      const { u1, u2 } = await this.readState();
      const local = index.map((x, a) => x - 16 * c[a]);
      const cell = 16 ** this.ndim * i + local[0] + 16 * (local[1] + 16 * local[2]);
      const dx = this.dx0 / (1 << L);
      const lo = domainLo({ ...this.cfg, dx: this.dx0, n: level0(this.root, this.ndim) });
      return {
        // (a 2D block's cells lie on z = 0, blk_pos, src/amr/grid.wgsl:172-177)
        pos: index.map((x, a) => (a === 2 && this.ndim === 2 ? 0 : (x + 0.5) * dx + lo[a])),
        index,
        dx,
        u1: Array.from(u1.subarray(4 * cell, 4 * cell + 4)),
        u2: Array.from(u2.subarray(4 * cell, 4 * cell + 4)),
        block: { level: L, coords: c, slot: i, indicator: null },
        stencil: undefined,
      };
    }
    return null;
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

  /** Per leaf the refinement indicator (amr/flag.wgsl's Loehner estimate, largest over its cells); NaN for parents. */
  indicator() {
    return native.amrIndicator(live(this.h));
  }

  /**
   * Re-mesh to `blocks` (Int32Array rows), within one level of the current tree everywhere. The
   * state reset() loads was of the old tree: it is dropped.
   */
  regrid(blocks) {
    native.amrRegrid(live(this.h), blocks);
    this.tree(blocks);
    this.initial = null;
  }

  /** AmrSim.runToEnd (src/amr/amr_sim.ts:3749-3775): the clock's step count once it holds at tEnd. */
  async runToEnd(batch = 16, maxSteps = 1e6) {
    if (!Number.isFinite(this.cfg.tEnd)) throw new Error('runToEnd needs a finite tEnd');
    const n = native.runToEnd(live(this.h), batch, Math.min(maxSteps, 0x7fffffff));
    if (this.regridEvery > 0) this.syncTree();
    return n;
  }

  /** AmrSim.writeParams (src/amr/amr_sim.ts:1677): this.cfg's run-time fields, from the next step on. */
  writeParams() {
    native.writeParams(live(this.h), this.nativeConfig());
  }

  /** AmrSim.buildHydro (src/amr/amr_sim.ts:1107): the block kernels of this.cfg's scheme. */
  async buildHydro() {
    native.buildHydro(live(this.h), this.nativeConfig());
  }

  nativeConfig() {
    return toNative({ ...this.cfg, n: level0(this.root, this.ndim), ndim: this.ndim, dx: this.dx0 });
  }

  async profile(steps = 5) {
    const ms = native.amrProfile(live(this.h), steps);
    return { assemble: ms[0], clock: ms[1], predict: ms[2], update: ms[3], step: ms[4] };
  }
}

/** The level-0 grid's cells per axis: 16 per root block (one layer along z in 2D). */
function level0(root, ndim) {
  return [16 * root[0], 16 * root[1], ndim === 2 ? 1 : 16 * root[2]];
}

/** createSim (src/create.ts:58-61): a block grid when a tree is given, else the uniform grid. */
function createSim(cfg, tree) {
  return tree ? AmrSim.create(cfg, tree) : Sim.create(cfg);
}

module.exports = { configFor, createSim, Sim, AmrSim, diagnosticsFrom, hipUnsupported, toNative, version: native.version, native };
