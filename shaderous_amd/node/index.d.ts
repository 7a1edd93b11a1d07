// Types of the libshx Node binding: the reference's Simulation API for the hydro path
// (src/sim.ts SimConfig :48-264, Diagnostics :1810, Sim, src/amr/amr_sim.ts AmrSim).
export type Bc = 'periodic' | 'outflow' | 'reflect' | 'reflecting' | 'outflow-only' | 0 | 1 | 2 | 3;
export type Bc3 = [Bc, Bc, Bc];
export type State = Float32Array | Float64Array;

export interface SimConfig {
  n: number | [number, number, number];
  ndim: 1 | 2 | 3;
  box: number;
  dx?: number;
  bcLo: Bc3;
  bcHi: Bc3;
  riemann: 'hllc' | 'hll';
  limiter: 'koren' | 'mc' | 'vanleer' | 'minmod';
  isentropicRecon: boolean;
  eosParams: { gamma: number };
  cfl: number;
  tEnd: number;
  rhoFloor: number;
  pFloor: number;
  etaDual: number;
  shockThresh: number;
  fluff: [number, number, number];
  precision: 32 | 64;
  path: 'fused' | 'plain';
  device: number;
  tile: [number, number, number];
  /** the domain's lower corner (default -box/2 on every axis) */
  org?: [number, number, number];
  /** the state reset() loads */
  initialState?: { u1: State; u2: State };
  /** the clock a reset() starts from */
  restart?: { time: number; steps: number };
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

export interface AmrTree {
  root: [number, number, number];
  blocks: Int32Array;
  path?: 'fused' | 'passes' | 'levels';
  compensated?: boolean;
}

/** The members of the reference's `Simulation` (src/sim.ts:465-627) this path serves. */
interface Simulation {
  readonly ncell: number;
  readonly nz: number;
  readonly preparing: boolean;
  encodeSteps(enc: unknown, count: number, mayHold?: boolean, speculate?: number): number;
  submitted(): void;
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
  destroy(): void;
}

export declare class Sim implements Simulation {
  static create(cfg: SimConfig): Sim;
  readonly ncell: number;
  readonly nz: number;
  readonly preparing: boolean;
  encodeSteps(enc: unknown, count: number, mayHold?: boolean, speculate?: number): number;
  submitted(): void;
  stepsResume(): Promise<void> | null;
  step(count?: number): void;
  sync(): void;
  clock(): Float64Array;
  setState(u1: State, u2: State, rhoMean?: number): void;
  load(u1: State, u2: State, time?: number, steps?: number): void;
  reset(): void;
  readState(): Promise<{ u1: State; u2: State }>;
  readStateAsync(): Promise<{ u1: State; u2: State }>;
  diagnostics(): Promise<Diagnostics>;
  readDiagnostics(): Promise<Diagnostics | null>;
  runToEnd(batch?: number, maxSteps?: number): Promise<number>;
  writeParams(): void;
  buildHydro(): Promise<void>;
  fluxMismatch(): Promise<{ mismatched: number; faces: number }>;
  profile(steps?: number): Promise<{ rates: number; clock: number; predict: number; update: number; step: number }>;
  destroy(): void;
}

export declare class AmrSim implements Simulation {
  static create(cfg: SimConfig, tree: AmrTree): AmrSim;
  readonly nblocks: number;
  readonly ncell: number;
  readonly nz: number;
  readonly preparing: boolean;
  table(): Int32Array;
  stats(): { blocks: number; leaves: number; matchLeaves: number; levels: number; launchesPerStep: number };
  indicator(): Float64Array;
  regrid(blocks: Int32Array): void;
  encodeSteps(enc: unknown, count: number, mayHold?: boolean, speculate?: number): number;
  submitted(): void;
  stepsResume(): Promise<void> | null;
  step(count?: number): void;
  sync(): void;
  clock(): Float64Array;
  setState(u1: State, u2: State, rhoMean?: number): void;
  load(u1: State, u2: State, time?: number, steps?: number): void;
  reset(): void;
  readState(): Promise<{ u1: State; u2: State }>;
  readStateAsync(): Promise<{ u1: State; u2: State }>;
  diagnostics(): Promise<Diagnostics>;
  readDiagnostics(): Promise<Diagnostics | null>;
  profile(steps?: number): Promise<{ assemble: number; clock: number; predict: number; update: number; step: number }>;
  destroy(): void;
}

export declare function configFor(n: number | [number, number, number], over?: Partial<SimConfig>): SimConfig;
export declare function createSim(cfg: SimConfig): Sim;
export declare function createSim(cfg: SimConfig, tree: AmrTree): AmrSim;
export declare function version(): string;
