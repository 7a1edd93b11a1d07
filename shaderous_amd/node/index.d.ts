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
  scalar: number;
  ranges: [number, number][];
}

export interface AmrTree {
  root: [number, number, number];
  blocks: Int32Array;
  path?: 'fused' | 'passes' | 'levels';
  compensated?: boolean;
}

export declare class Sim {
  static create(cfg: SimConfig): Sim;
  readonly ncell: number;
  setState(u1: State, u2: State, time?: number, steps?: number): void;
  readState(): { u1: State; u2: State };
  readStateAsync(): Promise<{ u1: State; u2: State }>;
  step(count?: number): void;
  submitted(): void;
  runToEnd(batch?: number, maxSteps?: number): number;
  clock(): Float64Array;
  diagnostics(): Diagnostics;
  readDiagnostics(): Promise<Diagnostics>;
  fluxMismatch(): { mismatched: number; faces: number };
  profile(steps?: number): { rates: number; clock: number; predict: number; update: number; step: number };
  destroy(): void;
}

export declare class AmrSim {
  static create(cfg: SimConfig, tree: AmrTree): AmrSim;
  readonly nblocks: number;
  readonly ncell: number;
  table(): Int32Array;
  stats(): { blocks: number; leaves: number; matchLeaves: number; levels: number; launchesPerStep: number };
  indicator(): Float64Array;
  regrid(blocks: Int32Array): void;
  setState(u1: State, u2: State, time?: number, steps?: number): void;
  readState(): { u1: State; u2: State };
  readStateAsync(): Promise<{ u1: State; u2: State }>;
  step(count?: number): void;
  submitted(): void;
  clock(): Float64Array;
  diagnostics(): Diagnostics;
  profile(steps?: number): { assemble: number; clock: number; predict: number; update: number; step: number };
  destroy(): void;
}

export declare function configFor(n: number | [number, number, number], over?: Partial<SimConfig>): SimConfig;
export declare function createSim(cfg: SimConfig): Sim;
export declare function createSim(cfg: SimConfig, tree: AmrTree): AmrSim;
export declare function version(): string;
