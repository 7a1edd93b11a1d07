// Batched ensembles of independent 2D grids for gfx950: every instance of an ensemble (each its
// own size, boundaries and clock) advances in one launch per step that consumes a device list of
// column tiles drawn from all instances. The clock pass (hydro/clock.wgsl:132-242) of each
// instance, with that instance's largest rate, is taken inside the step (StepClock): every tile
// forms its instance's dt, the first workgroup writes every instance's next clock. Per step, either of
//   k_ens_fused   per tile (BX x BY columns, (BX - 4) x (BY - 4) cells), one plane: primitives
//                 into LDS, the Hancock half step of the cells and the one-column ring, each face
//                 once, the conservative update, and the next step's rate of the instance
//                 (4 barriers per tile)
// The arithmetic is the uniform grid's, in the order of its plain 2D passes (a flat grid's z
// neighbours are the cell itself: its half step leaves the zero z slope out, half_flat, as the
// plain predictor does on a flat grid), so an instance gives the values of a single 2D run of
// the plain path (zero fluxes added or left out change at most the sign of a zero).
#pragma once
#include "uniform_impl.hpp"

namespace shx {

// One tile of the work list: instance, and the tile's lower corner in cells
struct EnsTile {
  int inst, x0, y0, y1;   // y1: a strip's end row (k_ens_rows; the tiles ignore it)
};

// Per instance: its parameters, and where its cells start in the concatenated state
template <typename T>
struct EnsInst {
  Params<T> P;
  long long off;
};

// The largest rate of every instance's cells (the first step after a state change)
template <typename T>
__global__ __launch_bounds__(256) void k_ens_rates(const EnsInst<T>* __restrict__ inst, const EnsTile* __restrict__ tiles,
                                                   int ntiles, int tx, int ty, const T4<T>* __restrict__ U1,
                                                   const T4<T>* __restrict__ U2, unsigned long long* slot,
                                                   int stride) {
  const EnsTile w = tiles[blockIdx.x];
  const Params<T>& P = inst[w.inst].P;
  const long long off = inst[w.inst].off;
  T m = T(0);
  for (int k = threadIdx.x; k < tx * ty; k += 256) {
    const int i = w.x0 + k % tx, j = w.y0 + k / tx;
    if (i >= P.n[0] || j >= P.n[1]) continue;
    const size_t o = off + (size_t)j * P.n[0] + i;
    const T4<T> u2 = U2[o];
    if (u2.w <= T(0.5)) m = tmax(m, rate_of(P, prim_from_cons(P, U1[o], u2)));
  }
  block_max_to(m, slot + (size_t)stride * w.inst);
}

// LDS of the 2D tile kernel (structure of arrays over the NT columns): q (7), the upper x and y
// sides (9 each); the x and y fluxes of the face below each column (8 each) reuse the sides'
// space once every face has read its sides (a barrier apart)
template <typename T, int NT>
struct Ens2dLayout {
  static constexpr int Q = 0, SX = 7 * NT, SY = 16 * NT, FX = 7 * NT, FY = 15 * NT, SIZE = 25 * NT;
};

// SINGLE: a single grid (a flat uniform grid stepped as a batch of one) whose parameters come
// as the kernel argument P1 and whose tiles are raster (tile wi at (wi mod ntx, wi / ntx)), so a
// workgroup's first loads wait on neither the tile list nor the instance record. The clock pass
// of the next step is the last workgroup's, which the launch adds when every tile has one already.
template <typename T, int LIM, int RS, bool ISEN, int BX, int BY, bool SINGLE>
__global__ __launch_bounds__(BX* BY, (sizeof(T) == 4 ? 2 : 1)) void k_ens_fused(
    const EnsInst<T>* __restrict__ inst, const EnsTile* __restrict__ tiles, int ntiles, const T4<T>* __restrict__ U1,
    const T4<T>* __restrict__ U2, T4<T>* __restrict__ O1, T4<T>* __restrict__ O2, StepClock<T> C, int ninst,
    Params<T> P1, int ntx) {
  constexpr int NT = BX * BY;
  using LY = Ens2dLayout<T, NT>;
  __shared__ T L[LY::SIZE];
  __shared__ T red[NT / 64 + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int t = tx + BX * ty;
  const bool in_r1 = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2;
  if (blockIdx.x == gridDim.x - 1)
    for (int i = t; i < ninst; i += NT) step_clock_advance(SINGLE ? P1 : inst[i].P, C, i);

  for (int wi = blockIdx.x; wi < ntiles; wi += gridDim.x) {
    const EnsTile w = SINGLE ? EnsTile{0, (wi % ntx) * (BX - 4), (wi / ntx) * (BY - 4), 0} : tiles[wi];
    const Params<T> P = SINGLE ? P1 : inst[w.inst].P;
    const long long off = SINGLE ? 0ll : inst[w.inst].off;
    const int nx = P.n[0], ny = P.n[1];
    const T dt = step_dt(P, C, w.inst);
    const T dtdx = dt / P.dx;
    const int x = w.x0 + tx - 2, y = w.y0 + ty - 2;
    const AxisMap mx = map_axis(x, nx, P.bc_lo[0], P.bc_hi[0]);
    const AxisMap my = map_axis(y, ny, P.bc_lo[1], P.bc_hi[1]);
    const bool real_x = x >= 0 && x < nx, real_y = y >= 0 && y < ny;
    const bool cell = tx >= 2 && tx <= BX - 3 && ty >= 2 && ty <= BY - 3 && real_x && real_y;
    const bool xface = tx >= 2 && tx <= BX - 2 && ty >= 2 && ty <= BY - 3 && x <= nx && real_y;
    const bool yface = ty >= 2 && ty <= BY - 2 && tx >= 2 && tx <= BX - 3 && y <= ny && real_x;
    // a ghost column beyond a non-periodic face takes the half step of the cell it maps to
    // (load_pred, hydro/update.wgsl:47-63): that cell's stencil sits inward in the tile
    int sx = tx, sy = ty;
    if (!real_x && ((x < 0 && P.bc_lo[0] != BC_PERIODIC) || (x >= nx && P.bc_hi[0] != BC_PERIODIC))) sx = tx + (mx.i - x);
    if (!real_y && ((y < 0 && P.bc_lo[1] != BC_PERIODIC) || (y >= ny && P.bc_hi[1] != BC_PERIODIC))) sy = ty + (my.i - y);
    const bool mapped = (sx != tx) || (sy != ty);
    const int ts = min(max(sx, 1), BX - 2) + BX * min(max(sy, 1), BY - 2);

    // the previous tile's last LDS reads are done (a workgroup's first tile has none before it)
    if (wi != (int)blockIdx.x) __syncthreads();
    const size_t o = off + (size_t)my.i * nx + mx.i;
    const T4<T> u1 = U1[o], u2 = U2[o];
    Prim<T> q = prim_from_cons(P, u1, u2);
    q.vx = bc_vel(q.vx, mx.flip, mx.side);
    q.vy = bc_vel(q.vy, my.flip, my.side);
    lds_put_prim<T, NT>(L + LY::Q, t, q);
    __syncthreads();

    SideData<T> sxl, syl;
    T hp = T(0);
    if (in_r1) {
      const Prim<T> pc = lds_prim_c<T, NT>(L + LY::Q, t);
      const Prim<T> xm = lds_prim<T, NT>(L + LY::Q, t - 1), xp = lds_prim<T, NT>(L + LY::Q, t + 1);
      const Prim<T> ym = lds_prim<T, NT>(L + LY::Q, t - BX), yp = lds_prim<T, NT>(L + LY::Q, t + BX);
      Inc<T> ixu, ixd, iyu, iyd;
      Half<T> hc = half_flat<T, LIM, ISEN>(P, pc, xm, xp, ym, yp, dt, &ixu, &ixd, &iyu, &iyd);
      if (mapped) {
        // the mapped cell's own stencil: its centre is this column's primitives seen unflipped
        Prim<T> qc = pc;
        qc.vx = bc_vel(qc.vx, mx.flip, 0);
        qc.vy = bc_vel(qc.vy, my.flip, 0);
        const Prim<T> sxm = lds_prim<T, NT>(L + LY::Q, ts - 1), sxp = lds_prim<T, NT>(L + LY::Q, ts + 1);
        const Prim<T> sym = lds_prim<T, NT>(L + LY::Q, ts - BX), syp = lds_prim<T, NT>(L + LY::Q, ts + BX);
        hc = half_flat<T, LIM, ISEN>(P, qc, sxm, sxp, sym, syp, dt, nullptr, nullptr, nullptr, nullptr);
        hc.vx = bc_vel(hc.vx, mx.flip, mx.side);
        hc.vy = bc_vel(hc.vy, my.flip, my.side);
      }
      hp = hc.p;
      sxl = side_of(hc, ixd, T(-1));
      syl = side_of(hc, iyd, T(-1));
      lds_put_side<T, NT>(L, LY::SX, t, side_of(hc, ixu, T(1)));
      lds_put_side<T, NT>(L, LY::SY, t, side_of(hc, iyu, T(1)));
    }
    __syncthreads();

    Flux<T> Fx = zero_flux<T>(), Fy = zero_flux<T>();
    if (xface) Fx = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SX, t - 1), sxl, 0);
    if (yface) Fy = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SY, t - BX), syl, 1);
    __syncthreads();
    if (xface) lds_put_flux<T, NT>(L, LY::FX, t, Fx);
    if (yface) lds_put_flux<T, NT>(L, LY::FY, t, Fy);
    __syncthreads();

    T r = T(0);
    if (cell) {
      const Flux<T> xh = lds_get_flux<T, NT>(L, LY::FX, t + 1);
      const Flux<T> yh = lds_get_flux<T, NT>(L, LY::FY, t + BX);
      const Flux<T> net = net_of(Fx, xh, Fy, yh, zero_flux<T>(), zero_flux<T>());
      T4<T> o1, o2;
      advance(P, u1, u2, net, hp, dtdx, o1, o2);
      const size_t oc = off + (size_t)y * nx + x;
      O1[oc] = o1;
      O2[oc] = o2;
      r = rate_of(P, prim_from_cons(P, o1, o2));
    }
    // the instance's next-step rate: tile maximum, one atomicMax
    T v = wave_max(r);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
      T m = T(0);
      for (int k = 0; k < (NT + 63) / 64; k++) m = tmax(m, red[k]);
      max_slot(C.next(w.inst), (unsigned long long)order_bits(m));
    }
  }
}

// ---------------- the row march (FP32) ----------------
// One wave per strip of 60 cells x a run of rows: lane l owns column x0 - 2 + l (lanes 2-61 the
// cells, 0-1 and 62-63 the x halo) and marches up the rows; a column's x neighbours come from the
// lanes beside it (DPP wave shifts: no LDS, no barrier), its y neighbours from the registers of
// the rows below and above. The arithmetic is k_ens_fused's (the plain 2D passes' order), so the
// bits are the same. Ghosts beyond an outflow or reflecting face take the half step of the cell
// they map to: an x ghost from that cell's lane (a lane permutation), a ghost row below from the
// first row's half step one row later, a ghost row above from the last row's (outflow-only faces
// and axes periodic on one face only are refused when the ensemble is made).
constexpr int ENS_STRIP = 60;

// lane l receives lane l - 1's value (shr) or lane l + 1's (shl); the end lanes get 0
__device__ __forceinline__ float dpp_shr1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float dpp_shl1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_get(float v, int src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, v)));
}
template <bool LEFT>
__device__ __forceinline__ Prim<float> prim_beside(const Prim<float>& q) {
  auto f = [](float v) { return LEFT ? dpp_shr1(v) : dpp_shl1(v); };
  Prim<float> o;
  o.rho = f(q.rho); o.vx = f(q.vx); o.vy = f(q.vy); o.vz = f(q.vz); o.p = f(q.p); o.s = f(q.s);
  o.cs = 0.0f;   // as lds_prim: a neighbour's sound speed is not used
  return o;
}
__device__ __forceinline__ SideData<float> side_from_left(const SideData<float>& d) {
  SideData<float> o;
  o.rho = dpp_shr1(d.rho); o.vx = dpp_shr1(d.vx); o.vy = dpp_shr1(d.vy); o.vz = dpp_shr1(d.vz);
  o.p = dpp_shr1(d.p); o.s = dpp_shr1(d.s); o.hrho = dpp_shr1(d.hrho); o.hp = dpp_shr1(d.hp);
  o.shock = dpp_shr1(d.shock);
  return o;
}
__device__ __forceinline__ Flux<float> flux_from_right(const Flux<float>& F) {
  Flux<float> o;
  o.m = dpp_shl1(F.m); o.mx = dpp_shl1(F.mx); o.my = dpp_shl1(F.my); o.mz = dpp_shl1(F.mz);
  o.e = dpp_shl1(F.e); o.ei = dpp_shl1(F.ei); o.s = dpp_shl1(F.s); o.vn = dpp_shl1(F.vn);
  return o;
}
__device__ __forceinline__ Half<float> half_of_lane(const Half<float>& h, int src) {
  Half<float> o;
  o.rho = lane_get(h.rho, src); o.vx = lane_get(h.vx, src); o.vy = lane_get(h.vy, src); o.vz = lane_get(h.vz, src);
  o.p = lane_get(h.p, src); o.s = lane_get(h.s, src); o.shock = lane_get(h.shock, src); o.fo = lane_get(h.fo, src);
  return o;
}

// Work item: EnsTile{instance, x0, y0, y1} (rows [y0, y1)). Every lane runs every instruction
// (the DPP and permute reads need all 64 lanes); lane masks only guard the stores.
template <int LIM, int RS, bool ISEN>
__global__ __launch_bounds__(64) void k_ens_rows(const EnsInst<float>* __restrict__ inst,
                                                 const EnsTile* __restrict__ strips, int nstrips,
                                                 const float4* __restrict__ U1, const float4* __restrict__ U2,
                                                 float4* __restrict__ O1, float4* __restrict__ O2,
                                                 StepClock<float> C, int ninst) {
  using T = float;
  const EnsTile w = strips[blockIdx.x];
  const Params<T> P = inst[w.inst].P;
  const long long off = inst[w.inst].off;
  const int nx = P.n[0], ny = P.n[1];
  const T dt = step_dt(P, C, w.inst);
  const T dtdx = dt / P.dx;
  const int lane = threadIdx.x;
  if (blockIdx.x == 0)
    for (int i = lane; i < ninst; i += 64) step_clock_advance(inst[i].P, C, i);
  const int x = w.x0 - 2 + lane;
  const int y0 = w.y0, y1 = w.y1;
  const AxisMap mx = map_axis(x, nx, P.bc_lo[0], P.bc_hi[0]);
  const bool real_x = x >= 0 && x < nx;
  // an x ghost beyond an outflow / reflecting face: the half step of the cell it maps to, whose
  // lane is in this strip (that cell lies within two columns of the face)
  const bool mapped_x = !real_x && ((x < 0 && P.bc_lo[0] != BC_PERIODIC) || (x >= nx && P.bc_hi[0] != BC_PERIODIC));
  const int src = mapped_x ? mx.i - (w.x0 - 2) : lane;
  const bool cell_lane = lane >= 2 && lane < 2 + ENS_STRIP && real_x;
  // ghost rows beyond an outflow / reflecting face: below the first row, above the last
  const bool ymap_lo = y0 == 0 && P.bc_lo[1] != BC_PERIODIC;
  const bool ymap_hi = y1 == ny && P.bc_hi[1] != BC_PERIODIC;

  auto row = [&](int r) -> Prim<T> {
    const AxisMap my = map_axis(r, ny, P.bc_lo[1], P.bc_hi[1]);
    const size_t o = off + (size_t)my.i * nx + mx.i;
    Prim<T> q = prim_from_cons(P, U1[o], U2[o]);
    q.vx = bc_vel(q.vx, mx.flip, mx.side);
    q.vy = bc_vel(q.vy, my.flip, my.side);
    return q;
  };

  Prim<T> pm = row(y0 - 2), pc = row(y0 - 1);
  SideData<T> syu_prev;
  Inc<T> iyu_ghost;                 // the ghost row below's own y increment
  Half<T> hc_prev;                  // the row below's half step (for a ghost row above)
  Flux<T> fy_lo = zero_flux<T>(), axy_prev = zero_flux<T>();
  T hp_prev = T(0);
  T rate_max = T(0);

  for (int c = y0 - 1; c <= y1; c++) {
    const Prim<T> pp = row(c + 1);
    const Prim<T> xm = prim_beside<true>(pc), xp = prim_beside<false>(pc);
    Inc<T> ixu, ixd, iyu, iyd;
    Half<T> hc = half_flat<T, LIM, ISEN>(P, pc, xm, xp, pm, pp, dt, &ixu, &ixd, &iyu, &iyd);
    // a ghost row above: the last row's half step
    const bool ghost_hi = ymap_hi && c == ny;
    if (ghost_hi) {
      const AxisMap my = map_axis(c, ny, P.bc_lo[1], P.bc_hi[1]);
      hc = hc_prev;
      hc.vy = bc_vel(hc.vy, my.flip, my.side);
    }
    // an x ghost: its cell's half step (all lanes take part in the permute)
    {
      const Half<T> hm = half_of_lane(hc, src);
      if (mapped_x) {
        hc = hm;
        hc.vx = bc_vel(hc.vx, mx.flip, mx.side);
      }
    }
    const SideData<T> sxl = side_of(hc, ixd, T(-1));
    const SideData<T> syl = side_of(hc, iyd, T(-1));
    const SideData<T> syu = side_of(hc, iyu, T(1));
    // the x face on this lane's left (every lane, for the shifts)
    const Flux<T> Fx = face_from_sides<T, RS>(P, side_from_left(side_of(hc, ixu, T(1))), sxl, 0);
    const Flux<T> xh = flux_from_right(Fx);
    Flux<T> axy;
    axy.m = Fx.m - xh.m; axy.mx = Fx.mx - xh.mx; axy.my = Fx.my - xh.my; axy.mz = Fx.mz - xh.mz;
    axy.e = Fx.e - xh.e; axy.ei = Fx.ei - xh.ei; axy.s = Fx.s - xh.s; axy.vn = Fx.vn - xh.vn;
    if (c == y0 - 1 && ymap_lo) {
      // the ghost row below the first: its half step is the first row's, formed next
      iyu_ghost = iyu;
    } else {
      Flux<T> fy = zero_flux<T>();
      if (c >= y0) {
        SideData<T> lo = syu_prev;
        if (c == 0 && ymap_lo) {
          const AxisMap my = map_axis(-1, ny, P.bc_lo[1], P.bc_hi[1]);
          Half<T> hg = hc;
          hg.vy = bc_vel(hg.vy, my.flip, my.side);
          lo = side_of(hg, iyu_ghost, T(1));
        }
        fy = face_from_sides<T, RS>(P, lo, syl, 1);
      }
      // the update of row c - 1 (net_of's order: x, then y; the flat z adds zeros)
      if (c - 1 >= y0 && c - 1 < y1 && cell_lane) {
        Flux<T> net;
        net.m = axy_prev.m + (fy_lo.m - fy.m);
        net.mx = axy_prev.mx + (fy_lo.mx - fy.mx);
        net.my = axy_prev.my + (fy_lo.my - fy.my);
        net.mz = axy_prev.mz + (fy_lo.mz - fy.mz);
        net.e = axy_prev.e + (fy_lo.e - fy.e);
        net.ei = axy_prev.ei + (fy_lo.ei - fy.ei);
        net.s = axy_prev.s + (fy_lo.s - fy.s);
        net.vn = axy_prev.vn + (fy_lo.vn - fy.vn);
        const size_t o = off + (size_t)(c - 1) * nx + x;
        T4<T> o1, o2;
        advance(P, U1[o], U2[o], net, hp_prev, dtdx, o1, o2);
        O1[o] = o1;
        O2[o] = o2;
        rate_max = tmax(rate_max, rate_of(P, prim_from_cons(P, o1, o2)));
      }
      fy_lo = fy;
    }
    syu_prev = syu;
    axy_prev = axy;
    hp_prev = hc.p;
    hc_prev = hc;
    pm = pc;
    pc = pp;
  }
  const T v = wave_max(rate_max);
  if (lane == 0) max_slot(C.next(w.inst), (unsigned long long)order_bits(v));
}

}  // namespace shx

namespace shx {

// Work shapes of the ensemble: k_ens_fused's tiles (threads BX x BY, cells (BX - 4) x (BY - 4))
// and k_ens_rows' strips (60 cells x 8, 16 or 32 rows, FP32)
enum { ENS_20x20 = 0, ENS_32x16 = 1, ENS_32x32 = 2, ENS_ROWS16 = 3, ENS_ROWS32 = 4, ENS_ROWS8 = 5, ENS_NTILES = 6 };
inline void ens_tile_cells(int tile, int& tx, int& ty) {
  switch (tile) {
    case ENS_20x20: tx = 16; ty = 16; break;
    case ENS_32x16: tx = 28; ty = 12; break;
    case ENS_32x32: tx = 28; ty = 28; break;
    case ENS_ROWS16: tx = ENS_STRIP; ty = 16; break;
    case ENS_ROWS8: tx = ENS_STRIP; ty = 8; break;
    default: tx = ENS_STRIP; ty = 32; break;
  }
}

struct EnsOps {
  // the largest rate of each instance's state into slot[stride * instance]
  void (*rates)(hipStream_t, const void* inst, const void* tiles, int ntiles, int tile, const void* u1, const void* u2,
                unsigned long long* slot, int stride);
  // one step of every instance (count of them), its clock pass included; p1 (may be null): the
  // parameters of a single instance (count 1, its tiles in raster order), passed by value
  void (*fused)(hipStream_t, int tile, int grid, const void* inst, int count, const void* tiles, int ntiles,
                const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck, const void* p1);
};

template <typename T, int LIM, int RS, bool ISEN>
struct EnsImpl {
  static void rates(hipStream_t st, const void* inst, const void* tiles, int ntiles, int tile, const void* u1,
                    const void* u2, unsigned long long* slot, int stride) {
    int tx, ty;
    ens_tile_cells(tile, tx, ty);
    hipLaunchKernelGGL(k_ens_rates<T>, dim3(ntiles), dim3(256), 0, st, (const EnsInst<T>*)inst,
                       (const EnsTile*)tiles, ntiles, tx, ty, (const T4<T>*)u1, (const T4<T>*)u2, slot, stride);
  }
  template <int BX, int BY>
  static void fused_tile(hipStream_t st, int grid, const void* inst, int count, const void* tiles, int ntiles,
                         const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck, const void* p1) {
    // (single grid: raster tiles of (BX - 4) x (BY - 4) cells; ntx of them per row)
    Params<T> P1{};
    int ntx = 0;
    if (p1 && count == 1) {
      P1 = *static_cast<const Params<T>*>(p1);
      ntx = (P1.n[0] + BX - 5) / (BX - 4);
    }
    // every tile its own workgroup and one more for the next step's clock pass
    if (grid > ntiles) grid = ntiles + 1;
    if (ntx > 0)
      hipLaunchKernelGGL((k_ens_fused<T, LIM, RS, ISEN, BX, BY, true>), dim3(grid), dim3(BX, BY), 0, st,
                         (const EnsInst<T>*)inst, (const EnsTile*)tiles, ntiles, (const T4<T>*)u1, (const T4<T>*)u2,
                         (T4<T>*)o1, (T4<T>*)o2, step_clock<T>(ck), count, P1, ntx);
    else
      hipLaunchKernelGGL((k_ens_fused<T, LIM, RS, ISEN, BX, BY, false>), dim3(grid), dim3(BX, BY), 0, st,
                         (const EnsInst<T>*)inst, (const EnsTile*)tiles, ntiles, (const T4<T>*)u1, (const T4<T>*)u2,
                         (T4<T>*)o1, (T4<T>*)o2, step_clock<T>(ck), count, P1, ntx);
  }
  static void fused(hipStream_t st, int tile, int grid, const void* inst, int count, const void* tiles, int ntiles,
                    const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck, const void* p1) {
    if (grid < 1 || ntiles < 1) return;
    if constexpr (sizeof(T) == 8) {
      fused_tile<20, 20>(st, grid, inst, count, tiles, ntiles, u1, u2, o1, o2, ck, p1);
    } else {
      switch (tile) {
        case ENS_32x32: fused_tile<32, 32>(st, grid, inst, count, tiles, ntiles, u1, u2, o1, o2, ck, p1); break;
        case ENS_32x16: fused_tile<32, 16>(st, grid, inst, count, tiles, ntiles, u1, u2, o1, o2, ck, p1); break;
        case ENS_ROWS8:
        case ENS_ROWS16:
        case ENS_ROWS32:
          // one wave per strip, every strip its own workgroup
          hipLaunchKernelGGL((k_ens_rows<LIM, RS, ISEN>), dim3(ntiles), dim3(64), 0, st,
                             (const EnsInst<float>*)inst, (const EnsTile*)tiles, ntiles, (const float4*)u1,
                             (const float4*)u2, (float4*)o1, (float4*)o2, step_clock<float>(ck), count);
          break;
        default: fused_tile<20, 20>(st, grid, inst, count, tiles, ntiles, u1, u2, o1, o2, ck, p1); break;
      }
    }
  }
  static EnsOps ops() {
    EnsOps o;
    o.rates = rates;
    o.fused = fused;
    return o;
  }
};

EnsOps ens_ops_f32(int lim, int rs, int isen);
EnsOps ens_ops_other32(int lim, int rs, int isen);
EnsOps ens_ops_f64(int lim, int rs, int isen);

}  // namespace shx
