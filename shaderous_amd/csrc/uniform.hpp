// Interface between the C-ABI layer and the uniform-grid kernels: type-erased launchers, one
// table per (precision, limiter, Riemann solver, isentropic) combination.
#pragma once
#include <algorithm>
#include "diag.hpp"
#include "hydro.hpp"

namespace shx {

struct Launch {
  hipStream_t stream;
  const void* params;   // Params<T> of the sim's precision
  size_t ncell;
};

// Column tiles of the fused kernel: k_fused's (block BX x BY threads in raster order, tile
// (BX - 4) x (BY - 4) cells) and k_fused_r's (threads by role, tiles of 32 x 16 or 32 x 8 cells).
// FUSED_20x20 is a 16 x 16 cell tile: k_fused_r's in FP32, k_fused's (the only FP64 tile) in FP64.
enum { FUSED_20x20 = 0, FUSED_32x16 = 1, FUSED_32x32 = 2, FUSED_R32x16 = 3, FUSED_R32x8 = 4, FUSED_NTILES = 5 };
// cells per tile along x and y
inline void fused_tile_cells(int tile, int& tx, int& ty) {
  switch (tile) {
    case FUSED_20x20: tx = 16; ty = 16; break;
    case FUSED_32x16: tx = 28; ty = 12; break;
    case FUSED_32x32: tx = 28; ty = 28; break;
    case FUSED_R32x16: tx = 32; ty = 16; break;
    default: tx = 32; ty = 8; break;
  }
}

// The clocks of a fused step (the step kernel takes its own clock pass, StepClock in
// uniform_impl.hpp), per instance: two clocks of 16 values (parity 0, 1), three rate slots, two
// epochs; p: the parity of the state the step reads.
struct ClockRefs {
  void* clk;
  unsigned long long* slot;
  unsigned int* epoch;
  int p;
};
// Bytes of the clock block of n instances: the clocks (2 x 16 values of es bytes each), the
// slots, the epochs (offsets clock_slots_at, clock_epochs_at).
inline size_t clock_slots_at(int n, int es) { return (size_t)n * 32 * es; }
inline size_t clock_epochs_at(int n, int es) { return clock_slots_at(n, es) + (size_t)n * 3 * 8; }
inline size_t clock_block_bytes(int n, int es) { return clock_epochs_at(n, es) + (size_t)n * 2 * 4; }

struct UniformOps {
  void (*rates)(const Launch&, const void* u1, const void* u2, unsigned long long* slot);
  void (*clock)(const Launch&, void* clk, unsigned long long* slot);
  void (*predict)(const Launch&, const void* u1, const void* u2, void* pa, void* pb, const void* clk);
  void (*update)(const Launch&, const void* u1, const void* u2, const void* pa, const void* pb, void* o1, void* o2,
                 const void* clk, bool debug);
  // the fused step; za, zb: the half steps of planes 0 and nz - 1 (zplanes) when a z face is not
  // periodic, else null
  void (*fused)(const Launch&, int tile, const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck,
                const void* za, const void* zb, int zc, bool debug);
  void (*zplanes)(const Launch&, const void* u1, const void* u2, void* za, void* zb, const ClockRefs& ck);
  // the 24 diagnostics channels of the state into out (FP64), per-workgroup records in part
  void (*diag)(const Launch&, const DiagGeo&, const void* u1, const void* u2, double* part, int nblocks, double* out);
};

// Each translation unit instantiates a subset; a null table means "not built".
UniformOps uniform_ops_koren32(int rs, int isen);
UniformOps uniform_ops_mc32(int rs, int isen);
UniformOps uniform_ops_other32(int lim, int rs);
UniformOps uniform_ops_f64(int lim, int rs, int isen);

inline UniformOps uniform_ops(int precision, int lim, int rs, int isen) {
  if (precision == 64) return uniform_ops_f64(lim, rs, isen);
  if (lim == LIM_KOREN) return uniform_ops_koren32(rs, isen);
  if (lim == LIM_MC) return uniform_ops_mc32(rs, isen);
  if (isen) return UniformOps{};
  return uniform_ops_other32(lim, rs);
}

}  // namespace shx
