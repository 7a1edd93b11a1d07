// Types of the libshx Node binding: the reference's Simulation API for the hydro path
// (src/sim.ts SimConfig :48-264, Simulation :465-627, CellProbe :267-281, Diagnostics :1810,
// Sim, src/amr/amr_sim.ts AmrSim). Sim and AmrSim have every member of the reference's
// Simulation, typed so that either is assignable to it with no cast: members that hand out
// WebGPU objects are typed `never` (their getters throw), members of subsystems this path has
// none of return the reference's own "none" (null, false, 0, '') or throw.
export type Bc = 'periodic' | 'outflow' | 'reflect' | 'reflecting' | 'outflow-only' | number;
export type Bc3 = [Bc, Bc, Bc];
export type State = Float32Array | Float64Array;

/**
 * What the backend reads from a configuration. The reference's SimConfig (src/sim.ts:48-264)
 * has every required field (the box side as problem.box, fluff as { rho, u, v }); so has the
 * SimConfig this module's configFor() makes. Physics the backend does not run is refused at
 * creation when the reference's fields ask for it (gravity, other equations of state, cooling,
 * diffusion, radiation, implicit pressure, species, obstacles, non-Cartesian coordinates,
 * boundary codes above 3).
 */
export interface HydroConfig {
  n: number | [number, number, number];
  ndim: 1 | 2 | 3;
  box?: number;
  problem?: { box: number };
  dx?: number;
  bcLo: Bc3;
  bcHi: Bc3;
  /** 'hllc' or 'hll' */
  riemann: string;
  /** 'koren', 'mc', 'vanleer' or 'minmod' */
  limiter: string;
  isentropicRecon: boolean;
  eosParams: { gamma: number };
  cfl: number;
  tEnd: number;
  rhoFloor: number;
  pFloor: number;
  etaDual: number;
  shockThresh: number;
  fluff?: [number, number, number] | { rho: number; u: number; v: number; va?: number };
  /** per-block timesteps on blocks (SimConfig.timeBins) */
  timeBins?: number;
  /** levels of refinement below the root grid; > 0 turns on the block grid's re-mesh cadence */
  amrLevels?: number;
  /** the re-mesh's indicator thresholds (SimConfig.refineHi, refineLo) */
  refineHi?: number;
  refineLo?: number;
  precision?: 32 | 64;
  path?: 'fused' | 'plain';
  device?: number;
  tile?: [number, number, number];
  /** the domain's lower corner (default -box/2 on every axis) */
  org?: [number, number, number];
  /** the state reset() loads */
  initialState?: { u1: State; u2: State };
  /** the clock a reset() starts from */
  restart?: { time: number; steps: number };
}

/** configFor()'s configuration: every field set. */
export interface SimConfig extends HydroConfig {
  box: number;
  riemann: 'hllc' | 'hll';
  limiter: 'koren' | 'mc' | 'vanleer' | 'minmod';
  fluff: [number, number, number];
  precision: 32 | 64;
  path: 'fused' | 'plain';
  device: number;
  tile: [number, number, number];
}

export interface Diagnostics {
  t: number;
  dt: number;
  steps: number;
  held: boolean;
  mass: number;
  momentum: [number, number, number];
  eKin: number;
  eInt: number;
  eGrav: number;
  eTot: number;
  maxMach: number;
  floored: number;
  jeansUnresolved: number;
  nuclearHeating: number;
  com: [number, number, number];
  perturber: null;
  masses: [];
  /** (min, max) of rho, temperature, specific scalar, pressure and speed */
  ranges: [number, number][];
}

/** CellProbe (src/sim.ts:267-281) without the stencil. */
export interface CellProbe {
  pos: [number, number, number];
  index: number[];
  dx: number;
  u1: number[];
  u2: number[];
  block?: { level: number; coords: number[]; slot: number; indicator: null };
  stencil?: undefined;
}

export interface AmrTree {
  /** root blocks per axis; a 2D tree (cfg.ndim 2: 16^2-cell blocks, every iz 0) has root[2] = 1 */
  root: [number, number, number];
  blocks: Int32Array;
  /** default 'fused' in 3D; 'passes' in 2D and with timeBins (the fused march steps 3D blocks, one dt) */
  path?: 'fused' | 'passes' | 'levels';
  compensated?: boolean;
  /** per-block timesteps: cycles of 2^timeBins substeps (default cfg.timeBins; passes path) */
  timeBins?: number;
  /** re-mesh every N steps inside the steps (default 8 where cfg.amrLevels > 0, else 0) */
  regridEvery?: number;
  /** the deepest level a re-mesh refines to (default cfg.amrLevels) */
  maxLevel?: number;
  /** indicator thresholds of a re-mesh (default cfg.refineHi, cfg.refineLo: 0.8, 0.2) */
  refineAbove?: number;
  refineBelow?: number;
}

/** The reference's Simulation (src/sim.ts:465-627) as this backend serves it. */
interface Simulation<C extends HydroConfig> {
  /** WebGPU objects: the getters throw. */
  readonly device: never;
  readonly paramBuffer: never;
  readonly clockBuffer: never;
  readonly fieldShader: never;
  fieldResources(parity: number, speciesParity: number, radParity?: number): never;
  cfg: C;
  /** finest cells per side, and along z (1 in 2D) */
  readonly n: number;
  readonly nz: number;
  /** always 0: the native sim keeps its own parity */
  cur: number;
  speciesCur: number;
  readonly speciesStride: number;
  readonly ncell: number;
  readonly preparing: boolean;
  readonly displayDims: [number, number, number];
  /** Enqueue up to `count` steps on the sim's stream; the encoder is not used. Returns `count`. */
  encodeSteps(enc: unknown, count: number, mayHold?: boolean, speculate?: number): number;
  submitted(): void;
  /** Always null: no step waits on a refinement check here. */
  stepsResume(): Promise<void> | null;
  step(count?: number): void;
  sync(): void;
  clock(): Float64Array;
  setState(u1: State, u2: State, rhoMean?: number): void;
  /** The state and the clock it starts from (a restart). */
  load(u1: State, u2: State, time?: number, steps?: number): void;
  reset(): void;
  readState(): Promise<{ u1: State; u2: State }>;
  readStateAsync(): Promise<{ u1: State; u2: State }>;
  diagnostics(): Promise<Diagnostics>;
  readDiagnostics(): Promise<Diagnostics | null>;
  runToEnd(batch?: number, maxSteps?: number): Promise<number>;
  /** this.cfg's run-time fields (gamma, cfl, tEnd, floors, etaDual, shockThresh, fluff) from the next step on */
  writeParams(): void;
  /** the kernels of this.cfg's limiter, solver and reconstruction */
  buildHydro(): Promise<void>;
  probe(f: [number, number, number], opts?: { stencil?: boolean }): Promise<CellProbe | null>;
  /** No self-gravity, point masses, sources or debris on this path. */
  refreshPotential(): void;
  massChanged(mass: number): void;
  readonly massCount: 0;
  addMass(m: unknown): false;
  readMasses(): Promise<null>;
  removeAddedMasses(): void;
  setTracerFixed(index: number, fixed: boolean): void;
  removeMasses(indices: number[]): void;
  retuneMasses(): void;
  readGasLumps(): Promise<null>;
  /** An empty list is accepted; sources throw. */
  setSources(list: readonly unknown[]): void;
  debris(): Promise<null>;
  debrisLedger(): Promise<null>;
  status(): string;
  /** The brush and star placement throw. */
  encodeBrush(enc: unknown, op: unknown): never;
  placeStar(op: unknown): Promise<never>;
  settleStar(star: unknown): Promise<never>;
  destroy(): void;
}

export interface Sim<C extends HydroConfig = SimConfig> extends Simulation<C> {
  fluxMismatch(): Promise<{ mismatched: number; faces: number }>;
  profile(steps?: number): Promise<{ rates: number; clock: number; predict: number; update: number; step: number }>;
}
export declare const Sim: {
  new <C extends HydroConfig = SimConfig>(cfg: C): Sim<C>;
  create<C extends HydroConfig = SimConfig>(cfg: C): Sim<C>;
};

export interface AmrSim<C extends HydroConfig = SimConfig> extends Simulation<C> {
  readonly nblocks: number;
  readonly levels: number;
  table(): Int32Array;
  stats(): { blocks: number; leaves: number; matchLeaves: number; levels: number; launchesPerStep: number };
  indicator(): Float64Array;
  regrid(blocks: Int32Array): void;
  profile(steps?: number): Promise<{ assemble: number; clock: number; predict: number; update: number; step: number }>;
}
export declare const AmrSim: {
  new <C extends HydroConfig = SimConfig>(cfg: C, tree: AmrTree): AmrSim<C>;
  create<C extends HydroConfig = SimConfig>(cfg: C, tree: AmrTree): AmrSim<C>;
};

export declare function configFor(n: number | [number, number, number], over?: Partial<SimConfig>): SimConfig;
export declare function createSim<C extends HydroConfig>(cfg: C): Sim<C>;
export declare function createSim<C extends HydroConfig>(cfg: C, tree: AmrTree): AmrSim<C>;
export declare function version(): string;
