// The fused uniform-grid step with threads by role (k_fused_r): the same arithmetic as k_fused,
// with a different map of threads to the tile's columns and two barriers per plane instead of
// four. A tile is TX x TY cells (32 x 16, 32 x 8 or 16 x 16, so 256^3 and other power-of-two
// grids split into whole tiles); of its (TX + 4) x (TY + 4) columns, threads own, in this order:
//   [0, TX TY)               the cell columns, x fastest: whole waves of cell work
//   the next 2 TX + 2 TY     ring 1 without its corners (half steps): the top row and the right
//                            column first (they also solve the faces they share with the tile's
//                            cells), then the bottom row and the left column
//   the next 4               ring 1's corners (loads only: their primitives feed the transverse
//                            slopes of ring 1's ends)
//   the last 2 TX + 2 TY     ring 2 without its corners (loads only); ring 2's corners are never
//                            read
// The cell waves never carry halo lanes; at 16 x 16 the 64 half-step lanes of ring 1 are one
// wave. LDS holds the plane data in raster order of the columns, with the x/y fluxes apart from
// the sides (RoleLayout), so a plane needs one barrier after the half steps and one after the
// fluxes.
#pragma once

namespace shx {

template <int TX, int TY>
struct RoleTile {
  static constexpr int CX = TX + 4, CY = TY + 4, NC = CX * CY;   // LDS columns
  static constexpr int NCELL = TX * TY;
  static constexpr int NHALF = NCELL + 2 * TX + 2 * TY;          // threads that form half steps
  static constexpr int NFACE = NCELL + TX + TY;                  // threads that solve x/y faces
  static constexpr int NTH = NHALF + 4 + 2 * TX + 2 * TY;        // threads
  // waves per SIMD the launch bound asks for: one 12-wave workgroup per CU at 32 x 16 (168
  // VGPRs), two 7-wave workgroups otherwise (128 VGPRs)
  static constexpr int MINW = NCELL > 256 ? 3 : 4;
};

// LDS of k_fused_r, per column: primitives of two planes (rho, v, p, s: the centre's sound speed
// stays in a register), x and y sides, x and y fluxes: 46 values (132 KB at 32 x 16 cells, one
// workgroup per CU; 79.5 KB at 32 x 8, two per CU)
template <int NT>
struct RoleLayout {
  static constexpr int Q0 = 0, Q1 = 6 * NT, SX = 12 * NT, SY = 21 * NT, FX = 30 * NT, FY = 38 * NT;
  static constexpr int SIZE = 46 * NT;
};

template <typename T, int NT>
__device__ __forceinline__ void lds_put_prim6(T* L, int t, const Prim<T>& q) {
  L[0 * NT + t] = q.rho; L[1 * NT + t] = q.vx; L[2 * NT + t] = q.vy; L[3 * NT + t] = q.vz;
  L[4 * NT + t] = q.p; L[5 * NT + t] = q.s;
}

// thread -> column (cx, cy) in cell-tile coordinates, [-2, TX + 2) x [-2, TY + 2)
template <int TX, int TY>
__device__ __forceinline__ void role_column(int t, int& cx, int& cy) {
  using RT = RoleTile<TX, TY>;
  if (t < RT::NCELL) {
    if constexpr (TX == 16) {
      // a wave holds a 4-wide strip of 16 rows: with the LDS rows 20 columns apart its lanes
      // then reach distinct banks (4 rows of 16 would meet in a third of them)
      cx = 4 * (t >> 6) + (t & 3);
      cy = (t & 63) >> 2;
    } else {
      cx = t % TX;
      cy = t / TX;
    }
    return;
  }
  t -= RT::NCELL;
  if (t < TX) { cx = t; cy = TY; return; }             // ring 1: y faces on the top row
  t -= TX;
  if (t < TY) { cx = TX; cy = t; return; }             // x faces on the right column
  t -= TY;
  if (t < TX) { cx = t; cy = -1; return; }             // the bottom row
  t -= TX;
  if (t < TY) { cx = -1; cy = t; return; }             // the left column
  t -= TY;
  if (t < 4) { cx = (t & 1) ? TX : -1; cy = (t & 2) ? TY : -1; return; }   // the corners
  t -= 4;
  if (t < TX) { cx = t; cy = -2; return; }             // ring 2: bottom, top, left, right
  t -= TX;
  if (t < TX) { cx = t; cy = TY + 1; return; }
  t -= TX;
  if (t < TY) { cx = -2; cy = t; return; }
  t -= TY;
  cx = TX + 1; cy = t;
}

// Launch bound: 32 x 16, 708 threads (12 waves), one workgroup per CU at up to 168 VGPRs;
// 32 x 8 (420 threads) and 16 x 16 (388), 7 waves, two per CU at up to 128 VGPRs.
// z: periodic, or through a non-periodic face as k_fused (ZA, ZB: the boundary planes' half steps).
// Barriers: the half steps of plane c (sides, and plane c + 1's primitives) are visible after
// the first, its x/y fluxes after the second. The sides of plane c + 1 overwrite those of plane c
// after the second barrier of plane c (the last reads of sides were before it); the fluxes of
// plane c + 1 overwrite those of plane c after the first barrier of plane c + 1; each primitive
// buffer is rewritten a plane after its neighbours were last read from it.
template <typename T, int LIM, int RS, bool ISEN, int TX, int TY, bool DEBUG, bool MAPBC>
__global__ __launch_bounds__((RoleTile<TX, TY>::NTH), (RoleTile<TX, TY>::MINW)) void k_fused_r(
    Params<T> P, const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, T4<T>* __restrict__ O1,
    T4<T>* __restrict__ O2, StepClock<T> C, const T4<T>* __restrict__ ZA, const T4<T>* __restrict__ ZB, int zc) {
  using RT = RoleTile<TX, TY>;
  constexpr int NT = RT::NC, BX = RT::CX, BY = RT::CY;
  using LY = RoleLayout<NT>;
  __shared__ T L[LY::SIZE];
  __shared__ T red[NT / 64 + 1];
  const int tid = threadIdx.x;
  int cx, cy;
  role_column<TX, TY>(tid, cx, cy);
  const int t = (cy + 2) * BX + (cx + 2);          // raster index of the column
  const bool cell = tid < RT::NCELL;               // wave-uniform
  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  // XCD-aware tile order: the dispatcher deals workgroups to the 8 XCDs in turn (workgroup i to
  // XCD i mod 8), so XCD k takes the k-th of 8 contiguous runs of tiles (the first n mod 8 runs
  // one tile longer) and marches it in order, rows of tiles side by side: the halo columns two
  // tiles share are read through one L2. The same tiles are stepped either way.
  int bx, by, bz;
  {
    const int ntile = gridDim.x * gridDim.y * gridDim.z;
    const int lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int k = lin & 7, q = ntile >> 3, r = ntile & 7;
    const int m = k * q + min(k, r) + (lin >> 3);
    bx = m % gridDim.x;
    by = (m / gridDim.x) % gridDim.y;
    bz = m / (gridDim.x * gridDim.y);
  }
  const int x = bx * TX + cx;
  const int y = by * TY + cy;
  const int z0 = bz * zc;
  const int z1 = min(z0 + zc, nz);
  const T dt = step_dt(P, C, 0);
  const T dtdx = dt / P.dx;
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && tid == 0) step_clock_advance(P, C, 0);

  AxisMap mx = map_axis(x, nx, P.bc_lo[0], P.bc_hi[0]);
  AxisMap my = map_axis(y, ny, P.bc_lo[1], P.bc_hi[1]);
  const size_t col = (size_t)my.i * nx + mx.i;
  const size_t plane = (size_t)nx * ny;

  const bool in_r1 = tid < RT::NHALF;              // cells and ring 1 without its corners
  const bool ring_face = !cell && tid < RT::NFACE; // ring 1's top row and right column
  const bool real_x = x >= 0 && x < nx, real_y = y >= 0 && y < ny;
  const bool upd = cell && real_x && real_y;
  const bool xface = cx >= 0 && cx <= TX && cy >= 0 && cy < TY && x <= nx && real_y;
  const bool yface = cy >= 0 && cy <= TY && cx >= 0 && cx < TX && y <= ny && real_x;
  // a ghost column beyond a non-periodic face takes the half step of the cell it maps to
  // (load_pred, hydro/update.wgsl:47-63): that cell's stencil sits inward in the tile
  int sx = cx, sy = cy;
  if (!real_x && ((x < 0 && P.bc_lo[0] != BC_PERIODIC) || (x >= nx && P.bc_hi[0] != BC_PERIODIC))) sx = cx + (mx.i - x);
  if (!real_y && ((y < 0 && P.bc_lo[1] != BC_PERIODIC) || (y >= ny && P.bc_hi[1] != BC_PERIODIC))) sy = cy + (my.i - y);
  const bool mapped = MAPBC && ((sx != cx) || (sy != cy));
  const int ts = (min(max(sy, -1), TY) + 2) * BX + (min(max(sx, -1), TX) + 2);

  // z: periodic, or (ZA given) a non-periodic face that ghost planes map through
  const bool zbc = MAPBC && ZA != nullptr;
  auto plane_of = [&](int z) {
    const int zi = zbc ? map_axis(z, nz, P.bc_lo[2], P.bc_hi[2]).i : ((z % nz) + nz) % nz;
    return (size_t)zi * plane + col;
  };
  auto prim_of = [&](T4<T> u1, T4<T> u2, int z) {
    Prim<T> q = prim_from_cons(P, u1, u2);
    if constexpr (MAPBC) {
      // (x and y periodic: every column's map has flip 1 and side 0, which bc_vel leaves as is)
      q.vx = bc_vel(q.vx, mx.flip, mx.side);
      q.vy = bc_vel(q.vy, my.flip, my.side);
    }
    if (zbc && (z < 0 || z >= nz)) {
      const AxisMap mz = map_axis(z, nz, P.bc_lo[2], P.bc_hi[2]);
      q.vz = bc_vel(q.vz, mz.flip, mz.side);
    }
    return q;
  };

  T4<T> b1, b2;
  size_t o0 = plane_of(z0 - 2);
  lds_put_prim6<T, NT>(L + LY::Q0, t, prim_of(U1[o0], U2[o0], z0 - 2));
  o0 = plane_of(z0 - 1);
  T cs_c;   // this column's sound speed in plane c
  {
    Prim<T> q = prim_of(U1[o0], U2[o0], z0 - 1);
    lds_put_prim6<T, NT>(L + LY::Q1, t, q);
    cs_c = q.cs;
  }
  o0 = plane_of(z0);
  b1 = U1[o0]; b2 = U2[o0];
  __syncthreads();

  SideData<T> zs_prev;
  Flux<T> fz_lo = zero_flux<T>();
  Flux<T> axy_prev = zero_flux<T>();
  T rate_max = T(0);

  for (int c = z0 - 1; c <= z1; c++) {
    Prim<T> pp = prim_of(b1, b2, c + 1);
    if (c + 2 <= z1 + 1) {
      size_t o = plane_of(c + 2);
      b1 = U1[o];
      b2 = U2[o];
    }
    const bool odd = ((c - z0 + 1) & 1) != 0;
    const T* Lq = L + (odd ? LY::Q0 : LY::Q1);
    T* Lm = L + (odd ? LY::Q1 : LY::Q0);

    // 2. half step and sides of plane c (hydro/predictor.wgsl:46-209)
    SideData<T> sxl, syl, szl, szu;
    if (in_r1) {
      Prim<T> pc = lds_prim<T, NT>(Lq, t);
      pc.cs = cs_c;
      Prim<T> pm = lds_prim<T, NT>(Lm, t);
      Prim<T> xm = lds_prim<T, NT>(Lq, t - 1), xp = lds_prim<T, NT>(Lq, t + 1);
      Prim<T> ym = lds_prim<T, NT>(Lq, t - BX), yp = lds_prim<T, NT>(Lq, t + BX);
      Inc<T> ixu, ixd, iyu, iyd, izu, izd;
      AxisSlope<T> sl[3];
      sl[0] = axis_slope<T, LIM, ISEN>(P, xm, pc, xp, &ixu, &ixd);
      sl[1] = axis_slope<T, LIM, ISEN>(P, ym, pc, yp, &iyu, &iyd);
      sl[2] = axis_slope<T, LIM, ISEN>(P, pm, pc, pp, &izu, &izd);
      Half<T> hc;
      if (!MAPBC || !mapped) {
        T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy) + (pp.vz - pm.vz);
        T jump = tmax(jump_of(xm.p, xp.p), tmax(jump_of(ym.p, yp.p), jump_of(pm.p, pp.p)));
        hc = hancock(P, pc, sl, conv, jump, dt);
      } else {
        Prim<T> qc = pc, qm = pm, qp = pp;
        qc.vx = bc_vel(qc.vx, mx.flip, 0); qc.vy = bc_vel(qc.vy, my.flip, 0);
        qm.vx = bc_vel(qm.vx, mx.flip, 0); qm.vy = bc_vel(qm.vy, my.flip, 0);
        qp.vx = bc_vel(qp.vx, mx.flip, 0); qp.vy = bc_vel(qp.vy, my.flip, 0);
        Prim<T> sxm = lds_prim<T, NT>(Lq, ts - 1), sxp = lds_prim<T, NT>(Lq, ts + 1);
        Prim<T> sym = lds_prim<T, NT>(Lq, ts - BX), syp = lds_prim<T, NT>(Lq, ts + BX);
        AxisSlope<T> ms[3];
        ms[0] = axis_slope<T, LIM, ISEN>(P, sxm, qc, sxp, nullptr, nullptr);
        ms[1] = axis_slope<T, LIM, ISEN>(P, sym, qc, syp, nullptr, nullptr);
        ms[2] = axis_slope<T, LIM, ISEN>(P, qm, qc, qp, nullptr, nullptr);
        T conv = (sxp.vx - sxm.vx) + (syp.vy - sym.vy) + (qp.vz - qm.vz);
        T jump = tmax(jump_of(sxm.p, sxp.p), tmax(jump_of(sym.p, syp.p), jump_of(qm.p, qp.p)));
        hc = hancock(P, qc, ms, conv, jump, dt);
        hc.vx = bc_vel(hc.vx, mx.flip, mx.side);
        hc.vy = bc_vel(hc.vy, my.flip, my.side);
      }
      if (zbc && upd) zghost_half(P, ZA, ZB, c, (size_t)y * nx + x, hc);
      sxl = side_of(hc, ixd, T(-1));
      syl = side_of(hc, iyd, T(-1));
      szl = side_of(hc, izd, T(-1));
      szu = side_of(hc, izu, T(1));
      lds_put_side<T, NT>(L, LY::SX, t, side_of(hc, ixu, T(1)));
      lds_put_side<T, NT>(L, LY::SY, t, side_of(hc, iyu, T(1)));
    }
    lds_put_prim6<T, NT>(Lm, t, pp);
    cs_c = pp.cs;
    __syncthreads();

    // 3. the x and y faces below this column, and the z face below plane c
    const bool live = c >= z0 && c < z1;
    Flux<T> Fx = zero_flux<T>(), Fy = zero_flux<T>(), fz = zero_flux<T>();
    if (cell) {
      if (live && xface) Fx = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SX, t - 1), sxl, 0);
      if (live && yface) Fy = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SY, t - BX), syl, 1);
      if (upd && c >= z0) fz = face_from_sides<T, RS>(P, zs_prev, szl, 2);
      // 4. the x/y fluxes into LDS
      if (live && xface) lds_put_flux<T, NT>(L, LY::FX, t, Fx);
      if (live && yface) lds_put_flux<T, NT>(L, LY::FY, t, Fy);
    } else if (ring_face && live && (xface || yface)) {
      // ring 1: the right column's x face or the top row's y face, one solve along the lane's axis
      const bool ax = xface;
      const Flux<T> F = face_from_sides_rt<T, RS>(
          P, lds_get_side<T, NT>(L, ax ? LY::SX : LY::SY, ax ? t - 1 : t - BX), ax ? sxl : syl, ax ? 0 : 1);
      lds_put_flux<T, NT>(L, ax ? LY::FX : LY::FY, t, F);
    }
    __syncthreads();

    // 5. this plane's x/y net inflow, and the update of plane c - 1
    if (cell) {
      Flux<T> axy = zero_flux<T>();
      if (live && upd) {
        Flux<T> xh = lds_get_flux<T, NT>(L, LY::FX, t + 1);
        Flux<T> yh = lds_get_flux<T, NT>(L, LY::FY, t + BX);
        if constexpr (DEBUG) {
          axy.m = Fx.m; axy.mx = xh.m; axy.my = Fy.m; axy.mz = yh.m;
        } else {
          axy.m = (Fx.m - xh.m) + (Fy.m - yh.m);
          axy.mx = (Fx.mx - xh.mx) + (Fy.mx - yh.mx);
          axy.my = (Fx.my - xh.my) + (Fy.my - yh.my);
          axy.mz = (Fx.mz - xh.mz) + (Fy.mz - yh.mz);
          axy.e = (Fx.e - xh.e) + (Fy.e - yh.e);
          axy.ei = (Fx.ei - xh.ei) + (Fy.ei - yh.ei);
          axy.s = (Fx.s - xh.s) + (Fy.s - yh.s);
          axy.vn = (Fx.vn - xh.vn) + (Fy.vn - yh.vn);
        }
      }
      if (upd && c - 1 >= z0 && c - 1 < z1) {
        size_t o = (size_t)(c - 1) * plane + col;
        T4<T> o1, o2;
        if constexpr (DEBUG) {
          o1.x = axy_prev.m; o1.y = axy_prev.mx; o1.z = axy_prev.my; o1.w = axy_prev.mz;
          o2.x = fz_lo.m; o2.y = fz.m; o2.z = T(0); o2.w = T(0);
        } else {
          Flux<T> net;
          net.m = axy_prev.m + (fz_lo.m - fz.m);
          net.mx = axy_prev.mx + (fz_lo.mx - fz.mx);
          net.my = axy_prev.my + (fz_lo.my - fz.my);
          net.mz = axy_prev.mz + (fz_lo.mz - fz.mz);
          net.e = axy_prev.e + (fz_lo.e - fz.e);
          net.ei = axy_prev.ei + (fz_lo.ei - fz.ei);
          net.s = axy_prev.s + (fz_lo.s - fz.s);
          net.vn = axy_prev.vn + (fz_lo.vn - fz.vn);
          advance(P, U1[o], U2[o], net, zs_prev.hp, dtdx, o1, o2);
          Prim<T> w = prim_from_cons(P, o1, o2);
          rate_max = tmax(rate_max, rate_of(P, w));
        }
        O1[o] = o1;
        O2[o] = o2;
      }
      zs_prev = szu;
      fz_lo = fz;
      axy_prev = axy;
    }
  }
  if constexpr (!DEBUG) {
    T v = wave_max(rate_max);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      T m = T(0);
      for (int w = 0; w < (NT + 63) / 64; w++) m = tmax(m, red[w]);
      max_slot(C.next(0), (unsigned long long)order_bits(m));
    }
  }
}

}  // namespace shx
