// Interface between the C-ABI layer and the block-AMR kernels: type-erased launchers, one table
// per (precision, limiter, Riemann solver, isentropic) combination.
//
// Blocks are 16^3 cells with 2 ghost layers (20^3 texels), as the reference's block atlas
// (src/amr/grid.wgsl:7-9). Device data of a block grid of S slots:
//   tab      S x 32 ints: the block table (src/amr/tree.ts:42-53)
//   U1, U2   S x 4096 cells of 4 values, x fastest: the state of leaves (parents unused)
//   W1, W2   S x 8000 texels: primitives (rho, vx, vy, vz) and (p, c_s, s, 0) of every texel but
//            the edges and corners, which stay zero (the assembled W atlas)
//   PA, PB   S x 8000 texels: half-step states of leaves' cells and first ghost layer
//   lists    slot lists (all blocks, leaves, leaves without / with a refined neighbour), each
//            also split by level
#pragma once
#include "diag.hpp"
#include "hydro.hpp"

namespace shx {

constexpr int AMR_BS = 16;       // cells per block side
constexpr int AMR_NG = 2;        // ghost layers
constexpr int AMR_PAD = 20;      // texels per block side
constexpr int AMR_BT = 32;       // ints per block-table row
constexpr int AMR_MAXLEV = 8;
constexpr int AMR_MAX_BINS = 5;  // cycles of at most 32 substeps (MAX_TIME_BINS, amr_sim.ts:115)
constexpr int AMR_CELLS = AMR_BS * AMR_BS * AMR_BS;
constexpr int AMR_TEXELS = AMR_PAD * AMR_PAD * AMR_PAD;

// Block-table row (tree.ts writeRow :652-675)
enum { BT_LEVEL = 0, BT_X = 1, BT_FLAGS = 4, BT_PARENT = 5, BT_NB = 6, BT_CHILD = 12, BT_COARSE = 20 };

// Cell size per level, in the run's precision (rounded once from the FP64 value).
template <typename T>
struct Levels {
  T dx[AMR_MAXLEV];
};

// Per-block timesteps (src/amr/bins_on.wgsl): the cycle's bin of every slot and the substep
// being taken; a leaf whose step has not started on this substep is part way through a longer
// one, seen through its half-step prediction, which the persistent PA/PB atlases hold from the
// start of its step (amr/mid_on.wgsl). k null: every block steps every step (bins_off.wgsl).
struct AmrBins {
  const int* k = nullptr;
  int sub = 0;
  const void* pa = nullptr;
  const void* pb = nullptr;
};

struct AmrLaunch {
  hipStream_t stream;
  const void* params;   // Params<T>: n = level-0 cells, dx = level-0 cell size
  const void* levels;   // Levels<T>
  const int* tab;       // device block table
  double dx0;
  AmrBins bins;         // read by assemble, predict and update
};

// The fused march's queue and what its workgroups read besides the state: the queue of leaves
// (pairs of consecutive entries, -1 pads), per slot its shell, per (slot, face) its action (2 r:
// fluxes to register r on a coarser block; 2 c + 1: the means of the finer fluxes into
// coarse-face buffer c; -1 neither) and register index, per slot the finer leaves it waits for
// (CSR: dep_off, dep), the control words (qctl) and a done flag per slot.
enum { AMR_Q_NEXT = 0, AMR_Q_EPOCH = 1, AMR_Q_ERROR = 2, AMR_Q_WORDS = 4 };
// bound of one dependency wait, in sleeps of about 128 cycles (~0.2 s at 2.4 GHz)
constexpr int AMR_WAIT_SPINS = 1 << 22;
struct AmrQueue {
  const int* list;
  int count;
  int grid;              // workgroups (one per CU)
  const int* shell_of;
  const int* act;
  void* reg;
  void* cf;
  const int* reg_of;
  int comp;              // TwoSum-compensated means
  const int* dep_off;
  const int* dep;
  int* qctl;
  int* done;
};

struct AmrOps {
  void (*assemble)(const AmrLaunch&, const int* list, int count, const void* u1, const void* u2, void* w1, void* w2,
                   unsigned long long* slot);
  void (*clock)(const AmrLaunch&, void* clk, unsigned long long* slot);
  void (*predict)(const AmrLaunch&, const int* list, int count, const void* w1, const void* w2, void* pa, void* pb,
                  const void* clk);
  void (*update)(const AmrLaunch&, const int* list, int count, bool match, bool compensated, const void* u1,
                 const void* u2, const void* w1, const void* w2, const void* pa, const void* pb, void* o1, void* o2,
                 const void* clk, bool debug);
  void (*rates)(const AmrLaunch&, const int* list, int count, const void* u1, const void* u2,
                unsigned long long* slot);
  // the 24 diagnostics channels over the listed leaves into out (FP64), a record per leaf in part
  void (*diag)(const AmrLaunch&, const DiagGeo&, const int* list, int count, const void* u1, const void* u2,
               double* part, double* out);
  // the fused form: the ghost shells (per leaf shell index, the waves of work listed as (shell,
  // slot, face, chunk): work0 the ghost texels, work1 the half steps beside boundary images;
  // work2 the half steps of face-rim cells, one record of 8 ints per source cell); the march of
  // every leaf in one launch, from the queue Q
  void (*shell)(const AmrLaunch&, const int* work0, int n0, const int* work1, int n1, const int* work2, int n2,
                const void* u1, const void* u2, void* sh, const void* clk);
  void (*fused)(const AmrLaunch&, const AmrQueue& Q, const void* sh, const void* u1, const void* u2, void* o1, void* o2,
                const void* clk, unsigned long long* slot);
  // re-meshing: `count` new leaves listed as int4 (new slot, kind 1 copy / 2 prolong / 3
  // restrict, old slot, child octant), from the old tree (L: its table, params, stream) and
  // state into both buffers of the new state
  void (*regrid)(const AmrLaunch& L, const int* map, int count, const void* u1, const void* u2, void* n1a, void* n2a,
                 void* n1b, void* n2b);
  // the refinement indicator of the listed leaves into out[slot] (precision of the state)
  void (*indicator)(const AmrLaunch&, const int* list, int count, const void* u1, const void* u2, void* out);
  // per-block timesteps: each listed leaf's largest rate / dx into out[slot] (precision of the
  // state); the clock's bank of the substep just taken (bins_tick.wgsl MODE 0); the cycle's dt
  // from the reduced rate in slot (hydro/clock.wgsl with BANK off and CYCLE substeps); the
  // updated blocks copied back into the state (the commit)
  void (*block_rates)(const AmrLaunch&, const int* list, int count, const void* u1, const void* u2, void* out);
  void (*bank)(const AmrLaunch&, void* clk);
  void (*clock_cycle)(const AmrLaunch&, void* clk, unsigned long long* slot, int cycle);
  void (*commit)(const AmrLaunch&, const int* list, int count, const void* o1, const void* o2, void* u1, void* u2);
};

AmrOps amr_ops_f32(int lim, int rs, int isen);
AmrOps amr_ops_other32(int lim, int rs);
AmrOps amr_ops_koren_hll32(int isen);
AmrOps amr_ops_f64(int lim, int rs, int isen);
AmrOps amr2d_ops_f32(int lim, int rs, int isen);
AmrOps amr2d_ops_f64(int lim, int rs, int isen);

inline AmrOps amr_ops(int precision, int lim, int rs, int isen, int ndim = 3) {
  if (ndim == 2) return precision == 64 ? amr2d_ops_f64(lim, rs, isen) : amr2d_ops_f32(lim, rs, isen);
  return precision == 64 ? amr_ops_f64(lim, rs, isen) : amr_ops_f32(lim, rs, isen);
}

}  // namespace shx
