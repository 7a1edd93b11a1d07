// Block-AMR hydro kernels for gfx950, one kernel per reference pass (src/amr/amr_sim.ts
// encodeSteps :2850-2929):
//   k_amr_assemble  amr/assemble.wgsl:77-154: every texel of every listed block, ghosts computed
//                   on load (same-level copy, boundary image, prolongation from a coarser leaf,
//                   restriction of finer children) -> W; the largest rate / dx of leaf cells
//   k_clock         hydro/clock.wgsl with PER_DX (dx taken as 1: the rates are per unit dx)
//   k_amr_predict   amr/predict.wgsl:183-221: half steps of leaves' cells and first ghost layer
//   k_amr_update    amr/update.wgsl:326-560 with the face tile (amr/tile.wgsl:59-104) and, for
//                   leaves with a refined neighbour, match_faces (amr/update.wgsl:207-290): a
//                   face on a refined neighbour takes the mean of the four fine fluxes,
//                   optionally summed with TwoSum (error-free FP32 addition) before the mean
// Each kernel takes a slot list: the host launches them once per level (the plain form) or once
// over every level (batched); the results are the same bits either way. D is the grid's
// dimension: 3, or 2 for quadtrees of 16^2-cell blocks (the is2d() branches of the passes: the
// z index is 0 throughout, no z neighbours, four children per parent, two fine faces per coarse
// face).
#pragma once
#include "amr.hpp"
#include "uniform_impl.hpp"

namespace shx {

__device__ __forceinline__ int bt(const int* __restrict__ tab, int s, int k) { return tab[s * AMR_BT + k]; }
__device__ __forceinline__ bool is_leaf(const int* __restrict__ tab, int s) { return (bt(tab, s, BT_FLAGS) & 3) == 3; }
// texel (x, y, z), each in [-2, 18), of a block's 20^3 atlas region; cell (x, y, z) in [0, 16)
__device__ __forceinline__ int tex(int x, int y, int z) {
  return ((z + AMR_NG) * AMR_PAD + (y + AMR_NG)) * AMR_PAD + (x + AMR_NG);
}
__device__ __forceinline__ int cel(int x, int y, int z) { return (z * AMR_BS + y) * AMR_BS + x; }
// per block: cells, texels (20^D) and the predicted texels (cells and first ghost layer, 18^D)
template <int D>
struct AmrGeo {
  static constexpr int CELLS = D == 3 ? AMR_CELLS : AMR_BS * AMR_BS;
  static constexpr int TEXELS = D == 3 ? AMR_TEXELS : AMR_PAD * AMR_PAD;
  static constexpr int ZC = D == 3 ? AMR_BS : 1;   // cells along z
};
template <int D>
__device__ __forceinline__ int texd(int x, int y, int z) {
  if constexpr (D == 3) return tex(x, y, z);
  return (y + AMR_NG) * AMR_PAD + (x + AMR_NG);
}
template <int D>
__device__ __forceinline__ int celd(int x, int y, int z) {
  if constexpr (D == 3) return cel(x, y, z);
  return y * AMR_BS + x;
}
// The cell size of level L: the level-0 size times 2^-L, exact (the same value as rounding
// dx0 / 2^L once); formed arithmetically, as an indexed kernel-argument array would be copied
// to scratch memory
template <typename T>
__device__ __forceinline__ T level_dx(const Params<T>& P, int L) {
  return P.dx * (T(1) / T(1 << L));
}

// ghost_src (amr/grid.wgsl:184-222): where texel (x, y, z) of block `slot` takes its state.
// kind 0: cell (x, y, z) of block `slot` (itself, a same-level neighbour or a refined one, or the
// boundary image in itself, with velocity flips and outflow-only sides); kind 1: interpolated
// from the coarser leaf `slot`; kind 2: an edge or corner texel, never assembled.
struct GSrc {
  int kind, slot, x, y, z;
  int fx, fy, fz;   // velocity sign per axis
  int sx, sy, sz;   // outflow-only side per axis
};
__device__ __forceinline__ GSrc ghost_src(const int* __restrict__ tab, int slot, int x, int y, int z) {
  GSrc g{0, slot, x, y, z, 1, 1, 1, 0, 0, 0};
  const bool ox = x < 0 || x >= AMR_BS, oy = y < 0 || y >= AMR_BS, oz = z < 0 || z >= AMR_BS;
  const int count = (int)ox + (int)oy + (int)oz;
  if (count == 0) return g;
  if (count > 1) { g.kind = 2; return g; }
  const int axis = ox ? 0 : (oy ? 1 : 2);
  int li = ox ? x : (oy ? y : z);
  const int upper = li >= AMR_BS;
  const int nb = bt(tab, slot, BT_NB + 2 * axis + upper);
  int flip = 1, side = 0;
  if (nb >= 0) {
    g.slot = nb;
    li += upper ? -AMR_BS : AMR_BS;
  } else if (nb == -1) {
    g.kind = 1;
    g.slot = bt(tab, slot, BT_COARSE + 2 * axis + upper);
    return g;
  } else {
    const int k = -1 - nb;
    if (k == BC_REFLECT) {
      li = upper ? 2 * AMR_BS - 1 - li : -1 - li;
      flip = -1;
    } else {
      li = upper ? AMR_BS - 1 : 0;
    }
    if (k == BC_OUTFLOW_ONLY) side = upper ? 1 : -1;
  }
  if (axis == 0) { g.x = li; g.fx = flip; g.sx = side; }
  else if (axis == 1) { g.y = li; g.fy = flip; g.sy = side; }
  else { g.z = li; g.fz = flip; g.sz = side; }
  return g;
}

template <typename T>
struct Cons {
  T4<T> a, b;
};
template <typename T>
__device__ __forceinline__ T4<T> v4(T x, T y, T z, T w) {
  T4<T> r;
  r.x = x; r.y = y; r.z = z; r.w = w;
  return r;
}

template <typename T, int D = 3>
__device__ __forceinline__ Cons<T> u_leaf(const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, int slot, int x,
                                          int y, int z) {
  size_t o = (size_t)slot * AmrGeo<D>::CELLS + celd<D>(x, y, z);
  return Cons<T>{U1[o], U2[o]};
}

// u_at (amr/cons.wgsl:19-43): a leaf's cell, or a parent's cell as the mean of its 8 children's
// (4 in 2D, each of weight kids_w(), amr/grid.wgsl:51-53)
template <typename T, int D = 3>
__device__ __forceinline__ Cons<T> u_at(const int* __restrict__ tab, const T4<T>* __restrict__ U1,
                                        const T4<T>* __restrict__ U2, int slot, int x, int y, int z) {
  if (is_leaf(tab, slot)) return u_leaf<T, D>(U1, U2, slot, x, y, z);
  const int hx = x >= AMR_BS / 2, hy = y >= AMR_BS / 2, hz = z >= AMR_BS / 2;
  const int ch = bt(tab, slot, BT_CHILD + (hx | (hy << 1) | (hz << 2)));
  if (ch < 0 || !is_leaf(tab, ch)) return Cons<T>{v4<T>(1, 0, 0, 0), v4<T>(1, 1, 0, 0)};
  const int fx = 2 * (x - (AMR_BS / 2) * hx), fy = 2 * (y - (AMR_BS / 2) * hy), fz = 2 * (z - (AMR_BS / 2) * hz);
  Cons<T> c{v4<T>(0, 0, 0, 0), v4<T>(0, 0, 0, 0)};
#pragma unroll
  for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int i = 0; i < 2; i++) {
        Cons<T> u = u_leaf<T, D>(U1, U2, ch, fx + i, fy + j, fz + k);
        c.a.x += u.a.x; c.a.y += u.a.y; c.a.z += u.a.z; c.a.w += u.a.w;
        c.b.x += u.b.x; c.b.y += u.b.y; c.b.z += u.b.z; c.b.w += u.b.w;
      }
  const T e = D == 3 ? T(0.125) : T(0.25);
  c.a.x *= e; c.a.y *= e; c.a.z *= e; c.a.w *= e;
  c.b.x *= e; c.b.y *= e; c.b.z *= e; c.b.w *= e;
  return c;
}

// u_near (amr/cons.wgsl:49-55): the state at a coarse block's texel, when it is a cell
template <typename T, int D = 3>
__device__ __forceinline__ bool u_near(const int* __restrict__ tab, const T4<T>* __restrict__ U1,
                                       const T4<T>* __restrict__ U2, int cs, int x, int y, int z, Cons<T>& out) {
  GSrc g = ghost_src(tab, cs, x, y, z);
  if (g.kind != 0) {
    out = Cons<T>{v4<T>(0, 0, 0, 0), v4<T>(0, 0, 0, 0)};
    return false;
  }
  Cons<T> u = u_at<T, D>(tab, U1, U2, g.slot, g.x, g.y, g.z);
  u.a.y = u.a.y * T(g.fx) + u.a.x * T(0);
  u.a.z = u.a.z * T(g.fy) + u.a.x * T(0);
  u.a.w = u.a.w * T(g.fz) + u.a.x * T(0);
  out = u;
  return true;
}

// Per-block timesteps: whether block s starts a step on the substep (blk_active,
// amr/bins_on.wgsl:14-17), and its own step (blk_dt, amr/blk_dt.wgsl)
__device__ __forceinline__ bool bins_active(const AmrBins& B, int s) {
  return !B.k || (B.sub & ((1 << B.k[s]) - 1)) == 0;
}
__device__ __forceinline__ int bins_bin(const AmrBins& B, int s) { return B.k ? B.k[s] : 0; }
// mid_block (amr/mid_on.wgsl:10-12): a leaf part way through a longer step
__device__ __forceinline__ bool mid_block(const AmrBins& B, const int* __restrict__ tab, int s) {
  return B.k && is_leaf(tab, s) && !bins_active(B, s);
}
// mid_cons (amr/mid_on.wgsl:14-30): the conserved state of texel (x, y, z) of block `slot` (a cell
// or a face ghost of the first layer) at the middle of its step, from its half-step prediction
template <typename T, int D = 3>
__device__ __forceinline__ Cons<T> mid_cons(const Params<T>& P, const AmrBins& B, int slot, int x, int y, int z) {
  const size_t o = (size_t)slot * AmrGeo<D>::TEXELS + texd<D>(x, y, z);
  const T4<T> a = static_cast<const T4<T>*>(B.pa)[o];
  const T4<T> b = static_cast<const T4<T>*>(B.pb)[o];
  const T rho = a.x;
  const T eint = b.x * P.igm1;
  Cons<T> c;
  c.a = v4<T>(rho, rho * a.y, rho * a.z, rho * a.w);
  c.b = v4<T>(eint + T(0.5) * rho * ((a.y * a.y + a.z * a.z) + a.w * a.w), eint, rho * b.y, T(0));
  return c;
}

// The limiter of the prolongation slopes (amr/cons.wgsl:65-74: MC for mc and koren)
template <int LIM>
constexpr int prolong_lim() { return (LIM == LIM_MINMOD || LIM == LIM_VANLEER) ? LIM : LIM_MC; }

// slope_between with mom_slope (amr/cons.wgsl:57-74): limited slopes of every conserved value,
// momentum by the product rule from the density and velocity slopes
template <typename T, int LIM>
__device__ __forceinline__ Cons<T> slope_between(const Params<T>& P, const Cons<T>& c, const Cons<T>& m,
                                                 const Cons<T>& p) {
  constexpr int PL = prolong_lim<LIM>();
  Cons<T> s;
  s.a.x = edge1<T, PL>(c.a.x - m.a.x, p.a.x - c.a.x, T(0));
  s.a.y = edge1<T, PL>(c.a.y - m.a.y, p.a.y - c.a.y, T(0));
  s.a.z = edge1<T, PL>(c.a.z - m.a.z, p.a.z - c.a.z, T(0));
  s.a.w = edge1<T, PL>(c.a.w - m.a.w, p.a.w - c.a.w, T(0));
  s.b.x = edge1<T, PL>(c.b.x - m.b.x, p.b.x - c.b.x, T(0));
  s.b.y = edge1<T, PL>(c.b.y - m.b.y, p.b.y - c.b.y, T(0));
  s.b.z = edge1<T, PL>(c.b.z - m.b.z, p.b.z - c.b.z, T(0));
  s.b.w = edge1<T, PL>(c.b.w - m.b.w, p.b.w - c.b.w, T(0));
  const T dc = tmax(c.a.x, P.rho_floor), dm = tmax(m.a.x, P.rho_floor), dp = tmax(p.a.x, P.rho_floor);
  const T vcx = tdiv(c.a.y, dc), vcy = tdiv(c.a.z, dc), vcz = tdiv(c.a.w, dc);
  const T svx = edge1<T, PL>(vcx - tdiv(m.a.y, dm), tdiv(p.a.y, dp) - vcx, T(0));
  const T svy = edge1<T, PL>(vcy - tdiv(m.a.z, dm), tdiv(p.a.z, dp) - vcy, T(0));
  const T svz = edge1<T, PL>(vcz - tdiv(m.a.w, dm), tdiv(p.a.w, dp) - vcz, T(0));
  s.a.y = vcx * s.a.x + c.a.x * svx;
  s.a.z = vcy * s.a.x + c.a.x * svy;
  s.a.w = vcz * s.a.x + c.a.x * svz;
  return s;
}

// prolong_slopes / prolong_child (amr/cons.wgsl:83-124): the fine child of parity (px, py, pz)
// of a coarse cell, from the 7-cell stencil (centre, -x, +x, -y, +y, -z, +z); one common
// scale-back keeps rho, E and e of every child at least half the coarse cell's
template <typename T, int LIM>
__device__ __forceinline__ Cons<T> prolong(const Params<T>& P, const Cons<T> v[7], const bool ok[7], int px, int py,
                                           int pz) {
  const Cons<T> z{v4<T>(0, 0, 0, 0), v4<T>(0, 0, 0, 0)};
  const Cons<T> sx = (ok[1] && ok[2]) ? slope_between<T, LIM>(P, v[0], v[1], v[2]) : z;
  const Cons<T> sy = (ok[3] && ok[4]) ? slope_between<T, LIM>(P, v[0], v[3], v[4]) : z;
  const Cons<T> sz = (ok[5] && ok[6]) ? slope_between<T, LIM>(P, v[0], v[5], v[6]) : z;
  const Cons<T>& c = v[0];
  T fit = T(1);
  {
    T d0 = ((T(0) + tabs(sx.a.x)) + tabs(sy.a.x)) + tabs(sz.a.x);
    T d1 = ((T(0) + tabs(sx.b.x)) + tabs(sy.b.x)) + tabs(sz.b.x);
    T d2 = ((T(0) + tabs(sx.b.y)) + tabs(sy.b.y)) + tabs(sz.b.y);
    T r0 = T(0.5) * tmax(c.a.x, T(0)), r1 = T(0.5) * tmax(c.b.x, T(0)), r2 = T(0.5) * tmax(c.b.y, T(0));
    fit = tmin(fit, tmin(T(1), tdiv(r0, tmax(T(0.25) * d0, T(1e-37)))));
    fit = tmin(fit, tmin(T(1), tdiv(r1, tmax(T(0.25) * d1, T(1e-37)))));
    fit = tmin(fit, tmin(T(1), tdiv(r2, tmax(T(0.25) * d2, T(1e-37)))));
  }
  const T ox = fit * (T(px) - T(0.5)) * T(0.5);
  const T oy = fit * (T(py) - T(0.5)) * T(0.5);
  const T oz = fit * (T(pz) - T(0.5)) * T(0.5);
  Cons<T> o;
#define SHX_PRO(F) o.F = ((c.F + ox * sx.F) + oy * sy.F) + oz * sz.F
  SHX_PRO(a.x); SHX_PRO(a.y); SHX_PRO(a.z); SHX_PRO(a.w);
  SHX_PRO(b.x); SHX_PRO(b.y); SHX_PRO(b.z); SHX_PRO(b.w);
#undef SHX_PRO
  return o;
}

// from_coarse (amr/assemble.wgsl:47-75): texel (x, y, z) of block `slot` interpolated from the
// coarser leaf cs. n0 = level-0 cells per axis. In 2D the z neighbours are left unread (not ok).
template <typename T, int LIM, int D = 3>
__device__ __forceinline__ Cons<T> from_coarse(const Params<T>& P, const int* __restrict__ tab,
                                               const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, int slot,
                                               int x, int y, int z, int cs, const AmrBins& B = AmrBins{}) {
  const int lc = bt(tab, cs, BT_LEVEL);
  int c[3], par[3];
  const int l[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int gf = bt(tab, slot, BT_X + a) * AMR_BS + l[a];
    const int nc = P.n[a] << lc;
    int gc = gf >> 1;
    gc = ((gc % nc) + nc) % nc;
    c[a] = gc - bt(tab, cs, BT_X + a) * AMR_BS;
    par[a] = gf & 1;
  }
  if constexpr (D == 2) c[2] = 0;
  Cons<T> v[7];
  bool ok[7];
  if constexpr (D == 2) {
    v[5] = v[6] = Cons<T>{v4<T>(0, 0, 0, 0), v4<T>(0, 0, 0, 0)};
    ok[5] = ok[6] = false;
  }
  if (mid_block(B, tab, cs)) {
    // a coarser leaf half way through its step: its prediction, cells and first ghost layer
    // (amr/assemble.wgsl:60-68)
    v[0] = mid_cons<T, D>(P, B, cs, c[0], c[1], c[2]);
    v[1] = mid_cons<T, D>(P, B, cs, c[0] - 1, c[1], c[2]);
    v[2] = mid_cons<T, D>(P, B, cs, c[0] + 1, c[1], c[2]);
    v[3] = mid_cons<T, D>(P, B, cs, c[0], c[1] - 1, c[2]);
    v[4] = mid_cons<T, D>(P, B, cs, c[0], c[1] + 1, c[2]);
#pragma unroll
    for (int k = 0; k < 5; k++) ok[k] = true;
    if constexpr (D == 3) {
      v[5] = mid_cons<T, D>(P, B, cs, c[0], c[1], c[2] - 1);
      v[6] = mid_cons<T, D>(P, B, cs, c[0], c[1], c[2] + 1);
      ok[5] = ok[6] = true;
    }
    return prolong<T, LIM>(P, v, ok, par[0], par[1], par[2]);
  }
  ok[0] = u_near<T, D>(tab, U1, U2, cs, c[0], c[1], c[2], v[0]);
  ok[1] = u_near<T, D>(tab, U1, U2, cs, c[0] - 1, c[1], c[2], v[1]);
  ok[2] = u_near<T, D>(tab, U1, U2, cs, c[0] + 1, c[1], c[2], v[2]);
  ok[3] = u_near<T, D>(tab, U1, U2, cs, c[0], c[1] - 1, c[2], v[3]);
  ok[4] = u_near<T, D>(tab, U1, U2, cs, c[0], c[1] + 1, c[2], v[4]);
  if constexpr (D == 3) {
    ok[5] = u_near<T, D>(tab, U1, U2, cs, c[0], c[1], c[2] - 1, v[5]);
    ok[6] = u_near<T, D>(tab, U1, U2, cs, c[0], c[1], c[2] + 1, v[6]);
  }
  return prolong<T, LIM>(P, v, ok, par[0], par[1], par[2]);
}

// The assembled primitives of texel (x, y, z) of block `slot`; false for edges and corners.
template <typename T, int LIM, int D = 3>
__device__ __forceinline__ bool amr_texel(const Params<T>& P, const int* __restrict__ tab,
                                          const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, int slot, int x,
                                          int y, int z, Prim<T>& w, const AmrBins& B = AmrBins{}) {
  const GSrc g = ghost_src(tab, slot, x, y, z);
  if (g.kind == 2) return false;
  const Cons<T> u = g.kind == 0 ? (mid_block(B, tab, g.slot) ? mid_cons<T, D>(P, B, g.slot, g.x, g.y, g.z)
                                                             : u_at<T, D>(tab, U1, U2, g.slot, g.x, g.y, g.z))
                                : from_coarse<T, LIM, D>(P, tab, U1, U2, slot, x, y, z, g.slot, B);
  w = prim_from_cons(P, u.a, u.b);
  w.vx = bc_vel(w.vx, g.fx, g.sx);
  w.vy = bc_vel(w.vy, g.fy, g.sy);
  w.vz = bc_vel(w.vz, g.fz, g.sz);
  return true;
}

// ---------------- assemble ----------------
template <int D>
constexpr int asm_chunks() { return (AmrGeo<D>::TEXELS + 255) / 256; }   // 32 workgroups of 256 per block (2 in 2D)
constexpr int ASM_CHUNKS = asm_chunks<3>();

template <typename T, int LIM, int D = 3>
__global__ __launch_bounds__(256) void k_amr_assemble(Params<T> P, Levels<T> LV, const int* __restrict__ tab,
                                                      const int* __restrict__ list, const T4<T>* __restrict__ U1,
                                                      const T4<T>* __restrict__ U2, T4<T>* __restrict__ W1,
                                                      T4<T>* __restrict__ W2, unsigned long long* slot, AmrBins B,
                                                      int count) {
  constexpr int NCH = asm_chunks<D>();
  int b, t;
  if constexpr (D == 3) {
    b = blockIdx.x / NCH;
    t = (blockIdx.x % NCH) * 256 + threadIdx.x;
  } else if ((int)blockIdx.x < count) {
    // 2D: one workgroup per block for its 256 cells, then one per two blocks for their 128 face
    // ghosts each
    b = blockIdx.x;
    t = threadIdx.x;
  } else {
    b = 2 * ((int)blockIdx.x - count) + (int)(threadIdx.x >> 7);
    t = AMR_BS * AMR_BS + (int)(threadIdx.x & 127);
  }
  T r = T(0);
  int x, y, z;
  bool any = t < AmrGeo<D>::TEXELS && b < count;
  if constexpr (D == 3) {
    x = t % AMR_PAD - AMR_NG;
    y = (t / AMR_PAD) % AMR_PAD - AMR_NG;
    z = t / (AMR_PAD * AMR_PAD) - AMR_NG;
  } else {
    // 2D: the cells' workgroups take only cells, the others the face ghosts face by face (-x,
    // +x, -y, +y; two layers of 16), so no wave mixes the cells' copy with a ghost's
    // prolongation or restriction; the 16 corner texels are never assembled
    z = 0;
    if (t < AMR_BS * AMR_BS) {
      x = t % AMR_BS;
      y = t / AMR_BS;
    } else {
      const int g = t - AMR_BS * AMR_BS, f = g >> 5, layer = (g >> 4) & 1, i = g & 15;
      const int a = (f & 1) ? AMR_BS + layer : -1 - layer;
      x = f < 2 ? a : i;
      y = f < 2 ? i : a;
    }
  }
  if (any) {
    const int s = list[b];
    Prim<T> w;
    if (amr_texel<T, LIM, D>(P, tab, U1, U2, s, x, y, z, w, B)) {
      const size_t o = (size_t)s * AmrGeo<D>::TEXELS + texd<D>(x, y, z);
      W1[o] = v4<T>(w.rho, w.vx, w.vy, w.vz);
      W2[o] = v4<T>(w.p, w.cs, w.s, T(0));
      const bool inside = x >= 0 && y >= 0 && z >= 0 && x < AMR_BS && y < AMR_BS && z < AmrGeo<D>::ZC;
      if (inside && is_leaf(tab, s)) r = rate_of(P, w) / level_dx(P, bt(tab, s, BT_LEVEL));
    }
  }
  block_max_to(r, slot);
}

// The rates of the listed leaves' cells alone (the fused form's first step)
template <typename T, int D = 3>
__global__ __launch_bounds__(256) void k_amr_rates(Params<T> P, Levels<T> LV, const int* __restrict__ tab,
                                                   const int* __restrict__ list, const T4<T>* __restrict__ U1,
                                                   const T4<T>* __restrict__ U2, unsigned long long* slot) {
  const int s = list[blockIdx.x];
  const T dx = level_dx(P, bt(tab, s, BT_LEVEL));
  T r = T(0);
  for (int c = threadIdx.x; c < AmrGeo<D>::CELLS; c += 256) {
    const size_t o = (size_t)s * AmrGeo<D>::CELLS + c;
    r = tmax(r, rate_of(P, prim_from_cons(P, U1[o], U2[o])) / dx);
  }
  block_max_to(r, slot);
}

// Per-block timesteps: each listed leaf's largest rate / dx into out[slot] (the records
// amr/bins_init.wgsl reads), by the arithmetic of the assembly's rates
template <typename T, int D = 3>
__global__ __launch_bounds__(256) void k_amr_block_rates(Params<T> P, const int* __restrict__ tab,
                                                         const int* __restrict__ list, const T4<T>* __restrict__ U1,
                                                         const T4<T>* __restrict__ U2, T* __restrict__ out) {
  __shared__ T red[4];
  const int s = list[blockIdx.x];
  const T dx = level_dx(P, bt(tab, s, BT_LEVEL));
  T r = T(0);
  for (int c = threadIdx.x; c < AmrGeo<D>::CELLS; c += 256) {
    const size_t o = (size_t)s * AmrGeo<D>::CELLS + c;
    r = tmax(r, rate_of(P, prim_from_cons(P, U1[o], U2[o])) / dx);
  }
  r = wave_max(r);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) out[s] = tmax(tmax(red[0], red[1]), tmax(red[2], red[3]));
}

// bins_tick.wgsl MODE 0: bank the substep just taken (Kahan-summed t)
template <typename T>
__global__ void k_amr_bank(T* clk) {
  const T prev = clk[0];
  if (prev > T(0)) {
    const T y = prev - clk[2];
    const T t = clk[1] + y;
    const T d = t - clk[1];
    clk[2] = d - y;
    clk[1] = t;
    clk[3] = clk[3] + T(1);
  }
}

// hydro/clock.wgsl with PER_DX (dx 1), BANK off and CYCLE substeps (:119-131): the cycle's dt;
// clears the reduction slot
template <typename T>
__global__ void k_amr_clock_cycle(Params<T> P, T* clk, unsigned long long* slot, T cycle) {
  const T rate = from_bits<T>(*slot);
  T dt = P.cfl * T(1) / tmax(rate, T(1e-30));
  dt = tmin(dt, P.dt_max);
  const T remaining = P.t_end - clk[1];
  if (remaining <= T(1e-6) * tabs(P.t_end)) dt = T(0);
  else dt = tmin(dt, remaining / cycle);
  if (T(1e30) - clk[1] < T(0.5) * cycle * dt) dt = T(0);
  if (clk[10] != T(0)) dt = T(0);
  clk[0] = dt;
  *slot = 0ull;
}

// the commit: the listed blocks' updated cells back into the state
template <typename T, int D = 3>
__global__ __launch_bounds__(256) void k_amr_commit(const int* __restrict__ list, const T4<T>* __restrict__ O1,
                                                    const T4<T>* __restrict__ O2, T4<T>* __restrict__ U1,
                                                    T4<T>* __restrict__ U2) {
  constexpr int C = AmrGeo<D>::CELLS;
  const int s = list[blockIdx.x / (C / 256)];
  const size_t o = (size_t)s * C + (blockIdx.x % (C / 256)) * 256 + threadIdx.x;
  U1[o] = O1[o];
  U2[o] = O2[o];
}

// ---------------- predict ----------------
constexpr int PRED_SIDE = AMR_BS + 2;
template <int D>
constexpr int pred_texels() { return D == 3 ? PRED_SIDE * PRED_SIDE * PRED_SIDE : PRED_SIDE * PRED_SIDE; }
template <int D>
constexpr int pred_chunks() { return (pred_texels<D>() + 255) / 256; }   // 23 (2 in 2D)
constexpr int PRED_TEXELS = pred_texels<3>();
constexpr int PRED_CHUNKS = pred_chunks<3>();

template <typename T>
__device__ __forceinline__ Prim<T> load_w(const T4<T>* __restrict__ W1, const T4<T>* __restrict__ W2, size_t o) {
  const T4<T> a = W1[o], b = W2[o];
  Prim<T> q;
  q.rho = tabs(a.x); q.vx = a.y; q.vy = a.z; q.vz = a.w;
  q.p = b.x; q.cs = b.y; q.s = b.z;
  return q;
}

// predict_cell (amr/predict.wgsl:90-181) on the atlas of block `slot` at texel (x, y, z); a zero
// in mask drops that axis's slopes (ghosts next to a coarser block whose transverse neighbour
// is an unassembled edge texel). In 2D the cell stands for its own z neighbours (:103-108).
template <typename T, int LIM, bool ISEN, int D = 3>
__device__ __forceinline__ Half<T> predict_texel(const Params<T>& P, const T4<T>* __restrict__ W1,
                                                 const T4<T>* __restrict__ W2, int slot, int x, int y, int z, T mx,
                                                 T my, T mz, T dt) {
  const size_t base = (size_t)slot * AmrGeo<D>::TEXELS;
  const int c = texd<D>(x, y, z);
  const Prim<T> q = load_w<T>(W1, W2, base + c);
  const Prim<T> xm = load_w<T>(W1, W2, base + c - 1), xp = load_w<T>(W1, W2, base + c + 1);
  const Prim<T> ym = load_w<T>(W1, W2, base + c - AMR_PAD), yp = load_w<T>(W1, W2, base + c + AMR_PAD);
  const Prim<T> zm = D == 3 ? load_w<T>(W1, W2, base + c - AMR_PAD * AMR_PAD) : q;
  const Prim<T> zp = D == 3 ? load_w<T>(W1, W2, base + c + AMR_PAD * AMR_PAD) : q;
  AxisSlope<T> sl[3];
  sl[0] = axis_slope<T, LIM, ISEN>(P, xm, q, xp, nullptr, nullptr);
  sl[1] = axis_slope<T, LIM, ISEN>(P, ym, q, yp, nullptr, nullptr);
  sl[2] = axis_slope<T, LIM, ISEN>(P, zm, q, zp, nullptr, nullptr);
  const T mk[3] = {mx, my, mz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    Inc<T>& s = sl[a].s;
    s.rho *= mk[a]; s.vx *= mk[a]; s.vy *= mk[a]; s.vz *= mk[a]; s.p *= mk[a]; s.s *= mk[a];
  }
  const T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy) + (zp.vz - zm.vz);
  const T jump = tmax(jump_of(xm.p, xp.p), tmax(jump_of(ym.p, yp.p), jump_of(zm.p, zp.p)));
  return hancock(P, q, sl, conv, jump, dt);
}

template <typename T, int LIM, bool ISEN, int D = 3>
__global__ __launch_bounds__(256) void k_amr_predict(Params<T> P, Levels<T> LV, const int* __restrict__ tab,
                                                     const int* __restrict__ list, const T4<T>* __restrict__ W1,
                                                     const T4<T>* __restrict__ W2, T4<T>* __restrict__ PA,
                                                     T4<T>* __restrict__ PB, const T* __restrict__ clk, AmrBins B,
                                                     int count) {
  constexpr int NCH = pred_chunks<D>();
  int b, t;
  if constexpr (D == 3) {
    b = blockIdx.x / NCH;
    t = (blockIdx.x % NCH) * 256 + threadIdx.x;
  } else if ((int)blockIdx.x < count) {
    // 2D: one workgroup per leaf for its 256 cells, then one per four leaves for their 64
    // first-layer face ghosts each
    b = blockIdx.x;
    t = threadIdx.x;
  } else {
    b = 4 * ((int)blockIdx.x - count) + (int)(threadIdx.x >> 6);
    t = AMR_BS * AMR_BS + (int)(threadIdx.x & 63);
    if (b >= count) return;
  }
  if (t >= pred_texels<D>()) return;
  int x, y, z;
  if constexpr (D == 3) {
    x = t % PRED_SIDE - 1;
    y = (t / PRED_SIDE) % PRED_SIDE - 1;
    z = t / (PRED_SIDE * PRED_SIDE) - 1;
  } else {
    // 2D: the cells first (four whole waves), then the 64 first-layer face ghosts face by face;
    // the 4 corner texels are edges, never predicted
    z = 0;
    if (t < AMR_BS * AMR_BS) {
      x = t % AMR_BS;
      y = t / AMR_BS;
    } else {
      const int g = t - AMR_BS * AMR_BS, f = g >> 4, i = g & 15;
      if (g >= 64) return;
      const int a = (f & 1) ? AMR_BS : -1;
      x = f < 2 ? a : i;
      y = f < 2 ? i : a;
    }
  }
  const int s = list[b];
  const GSrc g = ghost_src(tab, s, x, y, z);
  if (g.kind == 2) return;
  // with one timestep a ghost whose source is a refined block (a parent, whose atlas the step
  // does not assemble) is not predicted: the faces on that block are its children's
  // (match_faces), and nothing else reads the ghost's half step
  if (!B.k && g.kind == 0 && g.slot != s && !is_leaf(tab, g.slot)) return;
  // with bins, a ghost from a leaf part way through a longer step was filled from its prediction
  // and is predicted in place (amr/predict.wgsl:193-197); the leaf's own step is 2^bin dt
  const bool stale = B.k && g.slot != s && mid_block(B, tab, g.slot);
  const bool at = g.kind == 0 && !stale;
  const T dt = B.k ? clk[0] * T(1 << bins_bin(B, s)) : clk[0];
  const int src = at ? g.slot : s;
  Params<T> Pl = P;
  Pl.dx = level_dx(P, bt(tab, src, BT_LEVEL));
  const T mx = (!at && (x == 0 || x == AMR_BS - 1)) ? T(0) : T(1);
  const T my = (!at && (y == 0 || y == AMR_BS - 1)) ? T(0) : T(1);
  const T mz = (!at && (z == 0 || z == AMR_BS - 1)) ? T(0) : T(1);
  Half<T> h = at ? predict_texel<T, LIM, ISEN, D>(Pl, W1, W2, src, g.x, g.y, g.z, T(1), T(1), T(1), dt)
                 : predict_texel<T, LIM, ISEN, D>(Pl, W1, W2, src, x, y, z, mx, my, mz, dt);
  if (at) {
    h.vx = bc_vel(h.vx, g.fx, g.sx);
    h.vy = bc_vel(h.vy, g.fy, g.sy);
    h.vz = bc_vel(h.vz, g.fz, g.sz);
  }
  const size_t o = (size_t)s * AmrGeo<D>::TEXELS + texd<D>(x, y, z);
  PA[o] = v4<T>(h.rho, h.vx, h.vy, h.vz);
  PB[o] = v4<T>(h.p, h.s, h.shock, h.fo);
}

// ---------------- update ----------------
template <typename T>
__device__ __forceinline__ Half<T> load_half_at(const T4<T>* __restrict__ PA, const T4<T>* __restrict__ PB, size_t o) {
  const T4<T> a = PA[o], b = PB[o];
  Half<T> h;
  h.rho = a.x; h.vx = a.y; h.vy = a.z; h.vz = a.w;
  h.p = b.x; h.s = b.y; h.shock = b.z; h.fo = b.w;
  return h;
}

// flux_between on block s's atlas (hydro/flux.wgsl:254-347 with stored ghosts, amr/update.wgsl:
// 50-58): the face between texel (x, y, z) and the next one along axis
template <typename T, int LIM, int RS, bool ISEN, int D = 3>
__device__ __forceinline__ Flux<T> amr_flux(const Params<T>& P, const T4<T>* __restrict__ W1,
                                            const T4<T>* __restrict__ W2, const T4<T>* __restrict__ PA,
                                            const T4<T>* __restrict__ PB, int s, int x, int y, int z, int axis) {
  const size_t o = (size_t)s * AmrGeo<D>::TEXELS + texd<D>(x, y, z);
  const int e = axis == 0 ? 1 : (axis == 1 ? AMR_PAD : AMR_PAD * AMR_PAD);
  const Prim<T> wll = load_w<T>(W1, W2, o - e), wl = load_w<T>(W1, W2, o);
  const Prim<T> wr = load_w<T>(W1, W2, o + e), wrr = load_w<T>(W1, W2, o + 2 * e);
  const Half<T> hl = load_half_at<T>(PA, PB, o), hr = load_half_at<T>(PA, PB, o + e);
  const Inc<T> il = edge<T, LIM, ISEN, true>(P, wll, wl, wr);
  const Inc<T> ir = edge<T, LIM, ISEN, false>(P, wl, wr, wrr);
  return face_flux<T, RS>(P, hl, il, hr, ir, axis);
}

// TwoSum: s + e = a + b exactly (Knuth)
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
  const T x = a + b;
  const T bv = x - a;
  const T av = x - bv;
  e = (a - av) + (b - bv);
  s = x;
}

// The face of cell (x, y, z) of block s on side k (axis k / 2, upper if k odd), whose neighbour
// across it is refined: the mean of the four fine fluxes through it (match_faces,
// amr/update.wgsl:207-290; two in 2D, whose faces are edges: :246-251), summed in a fixed order,
// with TwoSum when COMP.
template <typename T, int LIM, int RS, bool ISEN, bool COMP, int D = 3>
__device__ __forceinline__ Flux<T> fine_face(const Params<T>& P, const int* __restrict__ tab,
                                             const T4<T>* __restrict__ W1, const T4<T>* __restrict__ W2,
                                             const T4<T>* __restrict__ PA, const T4<T>* __restrict__ PB, int np, int k,
                                             int x, int y, int z) {
  const int axis = k >> 1, hi = k & 1, upper = !hi;
  const int hx = x >= AMR_BS / 2, hy = y >= AMR_BS / 2, hz = z >= AMR_BS / 2;
  const int bx = axis == 0 ? upper : hx, by = axis == 1 ? upper : hy, bz = axis == 2 ? upper : hz;
  const int ch = bt(tab, np, BT_CHILD + (bx | (by << 1) | (bz << 2)));
  int ax = 2 * (x - (AMR_BS / 2) * hx), ay = 2 * (y - (AMR_BS / 2) * hy), az = 2 * (z - (AMR_BS / 2) * hz);
  const int an = upper ? AMR_BS - 1 : -1;
  if (axis == 0) ax = an; else if (axis == 1) ay = an; else az = an;
  const int t1 = axis == 2 ? 0 : axis + 1, t2 = axis == 0 ? 2 : axis - 1;   // (axis + 1) % 3, (axis + 2) % 3
  // fine faces along t1 and t2: 2 each, or 1 along z in 2D
  const int n1 = (D == 2 && t1 == 2) ? 1 : 2, n2 = (D == 2 && t2 == 2) ? 1 : 2;
  T sum[8], err[8];
#pragma unroll
  for (int q = 0; q < 8; q++) { sum[q] = T(0); err[q] = T(0); }
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      if (j >= n2 || i >= n1) continue;
      const int fx = ax + (t1 == 0 ? i : 0) + (t2 == 0 ? j : 0);
      const int fy = ay + (t1 == 1 ? i : 0) + (t2 == 1 ? j : 0);
      const int fz = az + (t1 == 2 ? i : 0) + (t2 == 2 ? j : 0);
      const Flux<T> f = amr_flux<T, LIM, RS, ISEN, D>(P, W1, W2, PA, PB, ch, fx, fy, fz, axis);
      const T fv[8] = {f.m, f.mx, f.my, f.mz, f.e, f.ei, f.s, f.vn};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if constexpr (COMP) {
          T e;
          two_sum(sum[q], fv[q], sum[q], e);
          err[q] = err[q] + e;
        } else {
          sum[q] = sum[q] + fv[q];
        }
      }
    }
  const T w = D == 3 ? T(0.25) : T(0.5);
  Flux<T> F;
  F.m = (sum[0] + err[0]) * w; F.mx = (sum[1] + err[1]) * w;
  F.my = (sum[2] + err[2]) * w; F.mz = (sum[3] + err[3]) * w;
  F.e = (sum[4] + err[4]) * w; F.ei = (sum[5] + err[5]) * w;
  F.s = (sum[6] + err[6]) * w; F.vn = (sum[7] + err[7]) * w;
  return F;
}

// The block update: workgroup 8 x 8 x 4 cells (a quarter-plane slab; 16 per block), every face
// of its box solved once into LDS (896 faces, 28 KB), then each cell's update. In 2D one
// workgroup of 16 x 16 cells per block (544 faces, 17 KB), no z faces.
constexpr int UPD_WX = 8, UPD_WY = 8, UPD_WZ = 4, UPD_TILES = 16;
template <int D>
struct UpdTile {
  static constexpr int WX = D == 3 ? UPD_WX : AMR_BS, WY = D == 3 ? UPD_WY : AMR_BS, WZ = D == 3 ? UPD_WZ : 1;
  static constexpr int TILES = D == 3 ? UPD_TILES : 1;
  static constexpr int NFX = (WX + 1) * WY * WZ, NFY = WX * (WY + 1) * WZ, NFZ = D == 3 ? WX * WY * (WZ + 1) : 0;
  static constexpr int NF = NFX + NFY + NFZ, NT = WX * WY * WZ;
};

// DEBUG (DEBUG_FLUX, amr/update.wgsl:324): instead of the update, each cell writes the mass
// flux of its six faces, (x lo, x hi, y lo, y hi) to O1 and (z lo, z hi, 0, 0) to O2.
template <typename T, int LIM, int RS, bool ISEN, bool MATCH, bool COMP, bool DEBUG = false, int D = 3>
__global__ __launch_bounds__(UpdTile<D>::NT) void k_amr_update(
    Params<T> P, Levels<T> LV, const int* __restrict__ tab, const int* __restrict__ list,
    const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, const T4<T>* __restrict__ W1,
    const T4<T>* __restrict__ W2, const T4<T>* __restrict__ PA, const T4<T>* __restrict__ PB, T4<T>* __restrict__ O1,
    T4<T>* __restrict__ O2, const T* __restrict__ clk, AmrBins B) {
  using UT = UpdTile<D>;
  constexpr int WX = UT::WX, WY = UT::WY, WZ = UT::WZ;
  constexpr int NFX = UT::NFX, NFY = UT::NFY;
  constexpr int NF = UT::NF, NT = UT::NT;
  __shared__ T tf[8 * NF];
  const int s = list[blockIdx.x / UT::TILES];
  const int tile = blockIdx.x % UT::TILES;
  const int x0 = (tile & 1) * WX, y0 = ((tile >> 1) & 1) * WY, z0 = (tile >> 2) * WZ;
  const int li = threadIdx.x;
  // per-block timesteps (amr/update.wgsl:173-186, 338-340): the block's bin and its own step; a
  // block half way through its step only deposits the second half of its faster neighbours'
  // fluxes (its own fluxes are zero)
  const int gbin = bins_bin(B, s);
  const bool deposit = B.k && !bins_active(B, s);
  if (!deposit)
  for (int f = li; f < NF; f += NT) {
    int axis, fx, fy, fz;
    if (f < NFX) {
      axis = 0; fx = f % (WX + 1); fy = (f / (WX + 1)) % WY; fz = f / ((WX + 1) * WY);
    } else if (f < NFX + NFY) {
      const int g = f - NFX;
      axis = 1; fx = g % WX; fy = (g / WX) % (WY + 1); fz = g / (WX * (WY + 1));
    } else {
      const int g = f - NFX - NFY;
      axis = 2; fx = g % WX; fy = (g / WX) % WY; fz = g / (WX * WY);
    }
    const Flux<T> F = amr_flux<T, LIM, RS, ISEN, D>(P, W1, W2, PA, PB, s, x0 + fx - (axis == 0),
                                                    y0 + fy - (axis == 1), z0 + fz - (axis == 2), axis);
    put_flux(tf, NF, f, F);
  }
  __syncthreads();
  const int tx = li % WX, ty = (li / WX) % WY, tz = li / (WX * WY);
  const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
  Flux<T> xl = zero_flux<T>(), xh = zero_flux<T>(), yl = zero_flux<T>(), yh = zero_flux<T>(), zl = zero_flux<T>(),
          zh = zero_flux<T>();
  if (!deposit) {
    xl = get_flux(tf, NF, tx + (WX + 1) * (ty + WY * tz));
    xh = get_flux(tf, NF, tx + 1 + (WX + 1) * (ty + WY * tz));
    yl = get_flux(tf, NF, NFX + tx + WX * (ty + (WY + 1) * tz));
    yh = get_flux(tf, NF, NFX + tx + WX * (ty + 1 + (WY + 1) * tz));
    if constexpr (D == 3) {
      zl = get_flux(tf, NF, NFX + NFY + tx + WX * (ty + WY * tz));
      zh = get_flux(tf, NF, NFX + NFY + tx + WX * (ty + WY * (tz + 1)));
    }
  }
  bool touched = false;
  if constexpr (MATCH) {
#pragma unroll
    for (int k = 0; k < 2 * D; k++) {
      const int axis = k >> 1, hi = k & 1;
      const int l = axis == 0 ? x : (axis == 1 ? y : z);
      if (l != (hi ? AMR_BS - 1 : 0)) continue;
      const int np = bt(tab, s, BT_NB + k);
      if (np < 0) continue;
      Flux<T> F;
      if (!is_leaf(tab, np)) {
        // the refined neighbour's child on this face; with bins its share of this block's step
        // (share, amr/update.wgsl:182-186): one half if faster, none in a deposit if in step
        T part = T(1);
        if (B.k) {
          const int hx = x >= AMR_BS / 2, hy = y >= AMR_BS / 2, hz = z >= AMR_BS / 2;
          const int upper = !hi;
          const int bx = axis == 0 ? upper : hx, by = axis == 1 ? upper : hy, bz = axis == 2 ? upper : hz;
          const int ch = bt(tab, np, BT_CHILD + (bx | (by << 1) | (bz << 2)));
          part = bins_bin(B, ch) < gbin ? T(0.5) : (deposit ? T(0) : T(1));
        }
        if (part == T(0)) {
          F = zero_flux<T>();
        } else {
          F = fine_face<T, LIM, RS, ISEN, COMP, D>(P, tab, W1, W2, PA, PB, np, k, x, y, z);
          if (B.k) {
            F.m = F.m * part; F.mx = F.mx * part; F.my = F.my * part; F.mz = F.mz * part;
            F.e = F.e * part; F.ei = F.ei * part; F.s = F.s * part; F.vn = F.vn * part;
          }
          touched = true;
        }
      } else if (B.k && np != s && bins_bin(B, np) < gbin) {
        // a faster leaf of the same level: half of its flux through the face, as it computes it
        const int an = hi ? -1 : AMR_BS - 1;
        F = amr_flux<T, LIM, RS, ISEN, D>(P, W1, W2, PA, PB, np, axis == 0 ? an : x, axis == 1 ? an : y,
                                          axis == 2 ? an : z, axis);
        F.m = F.m * T(0.5); F.mx = F.mx * T(0.5); F.my = F.my * T(0.5); F.mz = F.mz * T(0.5);
        F.e = F.e * T(0.5); F.ei = F.ei * T(0.5); F.s = F.s * T(0.5); F.vn = F.vn * T(0.5);
        touched = true;
      } else {
        continue;
      }
      switch (k) {
        case 0: xl = F; break;
        case 1: xh = F; break;
        case 2: yl = F; break;
        case 3: yh = F; break;
        case 4: zl = F; break;
        default: zh = F; break;
      }
    }
  }
  if constexpr (DEBUG) {
    const size_t o = (size_t)s * AmrGeo<D>::CELLS + celd<D>(x, y, z);
    O1[o] = v4<T>(xl.m, xh.m, yl.m, yh.m);
    O2[o] = v4<T>(zl.m, zh.m, T(0), T(0));
    return;
  }
  const size_t o = (size_t)s * AmrGeo<D>::CELLS + celd<D>(x, y, z);
  if (deposit && !touched) {
    // a cell on no face shared with a faster neighbour is copied as it is (amr/update.wgsl)
    O1[o] = U1[o];
    O2[o] = U2[o];
    return;
  }
  const Flux<T> net = net_of(xl, xh, yl, yh, zl, zh);
  const T dx = level_dx(P, bt(tab, s, BT_LEVEL));
  Params<T> Pl = P;
  Pl.dx = dx;
  const T dt = B.k ? clk[0] * T(1 << gbin) : clk[0];
  T4<T> o1, o2;
  advance(Pl, U1[o], U2[o], net, PB[(size_t)s * AmrGeo<D>::TEXELS + texd<D>(x, y, z)].x, dt / dx, o1, o2);
  O1[o] = o1;
  O2[o] = o2;
}

// ---------------- diagnostics (FP64, volume-weighted over leaves) ----------------
// The 24 channels of amr/diagnostics.wgsl:13-117: one workgroup per listed leaf (its cells of
// volume h^3 at blk_pos, amr/grid.wgsl:172-177), records folded by k_diag_final in list order.
template <typename T, int D = 3>
__global__ __launch_bounds__(256) void k_amr_diag(Params<T> P, DiagGeo G, const int* __restrict__ tab,
                                                  const int* __restrict__ list, double dx0,
                                                  const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2,
                                                  double* part) {
  __shared__ double sh[4 * DIAG_CH];
  const int s = list[blockIdx.x];
  const double h = dx0 / (double)(1 << bt(tab, s, BT_LEVEL));
  const double dv = G.flat ? h * h : h * h * h;
  const int bx = bt(tab, s, BT_X) * AMR_BS, by = bt(tab, s, BT_X + 1) * AMR_BS, bz = bt(tab, s, BT_X + 2) * AMR_BS;
  double r[DIAG_CH];
  diag_init(r);
  for (int c = threadIdx.x; c < AmrGeo<D>::CELLS; c += 256) {
    const int i = c % AMR_BS, j = (c / AMR_BS) % AMR_BS, k = c / (AMR_BS * AMR_BS);
    const size_t o = (size_t)s * AmrGeo<D>::CELLS + c;
    const double x = (bx + i + 0.5) * h + G.org[0], y = (by + j + 0.5) * h + G.org[1];
    const double z = G.flat ? 0.0 : (bz + k + 0.5) * h + G.org[2];
    diag_cell<T>(G, U1[o], U2[o], dv, x, y, z, r);
  }
  diag_block(r, sh, part + (size_t)blockIdx.x * DIAG_CH);
}

// ---------------- launchers ----------------
// ---------------- re-meshing ----------------
// Every leaf of a new tree from the old tree's state (amr/regrid.wgsl:40-139), per listed new
// leaf (new slot, kind, old source slot, child octant): 1 copy of an old leaf; 2 prolongation of
// an old leaf into its child (the ghost fill's conservative limited stencil over the old state,
// then the total energy as each child's kinetic energy plus its interpolated internal energy
// plus the offset that keeps the children's mean the coarse cell's, regrid.wgsl:86-97); 3
// restriction of an old parent (the mean of its leaf children, u_at). Written into both state
// buffers, as a state upload is. told: the old block table.
template <typename T>
__device__ __forceinline__ T ke_of(const Params<T>& P, const Cons<T>& u) {
  return tdiv(T(0.5) * ((u.a.y * u.a.y + u.a.z * u.a.z) + u.a.w * u.a.w), tmax(u.a.x, P.rho_floor));
}

template <typename T, int LIM, int D = 3>
__global__ __launch_bounds__(256) void k_amr_regrid(Params<T> P, const int* __restrict__ told,
                                                    const int4* __restrict__ map, const T4<T>* __restrict__ U1,
                                                    const T4<T>* __restrict__ U2, T4<T>* __restrict__ N1a,
                                                    T4<T>* __restrict__ N2a, T4<T>* __restrict__ N1b,
                                                    T4<T>* __restrict__ N2b) {
  constexpr int C = AmrGeo<D>::CELLS;
  const int4 m = map[blockIdx.x / (C / 256)];
  const int c = (blockIdx.x % (C / 256)) * 256 + threadIdx.x;
  const int x = c % AMR_BS, y = (c / AMR_BS) % AMR_BS, z = D == 3 ? c / (AMR_BS * AMR_BS) : 0;
  Cons<T> u;
  if (m.y == 1) {
    u = u_leaf<T, D>(U1, U2, m.z, x, y, z);
  } else if (m.y == 3) {
    u = u_at<T, D>(told, U1, U2, m.z, x, y, z);
  } else {
    const int lc[3] = {(x >> 1) + (AMR_BS / 2) * (m.w & 1), (y >> 1) + (AMR_BS / 2) * ((m.w >> 1) & 1),
                       D == 3 ? (z >> 1) + (AMR_BS / 2) * ((m.w >> 2) & 1) : 0};
    Cons<T> v[7];
    bool ok[7];
#pragma unroll
    for (int k = 0; k < 7; k++) {
      int at[3] = {lc[0], lc[1], lc[2]};
      if (k > 0) at[(k - 1) / 2] += (k & 1) == 0 ? 1 : -1;
      if (D == 2 && k >= 5) {
        // 2D: no z neighbours (regrid.wgsl:58-62)
        v[k] = Cons<T>{v4<T>(0, 0, 0, 0), v4<T>(0, 0, 0, 0)};
        ok[k] = false;
      } else {
        ok[k] = u_near<T, D>(told, U1, U2, m.z, at[0], at[1], at[2], v[k]);
      }
    }
    u = prolong<T, LIM>(P, v, ok, x & 1, y & 1, z & 1);
    T ke = T(0);
    for (int k = 0; k < (D == 3 ? 2 : 1); k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) ke = ke + ke_of(P, prolong<T, LIM>(P, v, ok, i, j, k));
    ke = ke * (D == 3 ? T(0.125) : T(0.25));
    u.b.x = ke_of(P, u) + u.b.y + ((v[0].b.x - v[0].b.y) - ke);
  }
  const size_t o = (size_t)m.x * C + c;
  N1a[o] = u.a; N2a[o] = u.b;
  N1b[o] = u.a; N2b[o] = u.b;
}

// The refinement indicator of the listed leaves (amr/flag.wgsl:66-155 for gas cells: the
// obstacle, species, burning, point-mass, density-threshold and Jeans terms belong to physics this
// path does not carry): per interior cell Loehner's normalized second derivative of density and
// of pressure,
//   chi^2 = sum_d (q+ - 2 q + q-)^2 / sum_d (|q+ - q| + |q - q-| + EPS (|q+| + 2 |q| + |q-|))^2,
// over the block's D axes, the neighbours taken from the assembled texels (ghosts included, as
// the flag pass reads the atlas the assembly left); the block's value is the largest
// max(chi_rho, chi_p) over its cells, at least 1e-30 (GAS_MIN).
constexpr double FLAG_EPS = 0.01, FLAG_GAS_MIN = 1e-30;

template <typename T>
__device__ __forceinline__ void flag_axis(T qm, T q0, T qp, T& num, T& den) {
  const T eps = T(FLAG_EPS);
  const T d = tabs(qp - q0) + tabs(q0 - qm) + eps * (tabs(qp) + T(2) * tabs(q0) + tabs(qm));
  const T s2 = qp - T(2) * q0 + qm;
  num = num + s2 * s2;
  den = den + d * d;
}

template <typename T, int LIM, int D = 3>
__global__ __launch_bounds__(256) void k_amr_indicator(Params<T> P, const int* __restrict__ tab,
                                                       const int* __restrict__ list, const T4<T>* __restrict__ U1,
                                                       const T4<T>* __restrict__ U2, T* __restrict__ out) {
  __shared__ T red[4];
  const int s = list[blockIdx.x];
  T r = T(0);
  for (int c = threadIdx.x; c < AmrGeo<D>::CELLS; c += 256) {
    const int x = c % AMR_BS, y = (c / AMR_BS) % AMR_BS, z = D == 3 ? c / (AMR_BS * AMR_BS) : 0;
    Prim<T> w0;
    amr_texel<T, LIM, D>(P, tab, U1, U2, s, x, y, z, w0);
    T nr = T(0), dr = T(0), np = T(0), dp = T(0);
#pragma unroll
    for (int a = 0; a < D; a++) {
      // a face neighbour of a cell is a cell or a face ghost: never an edge
      Prim<T> wm, wp;
      amr_texel<T, LIM, D>(P, tab, U1, U2, s, x - (a == 0), y - (a == 1), z - (a == 2), wm);
      amr_texel<T, LIM, D>(P, tab, U1, U2, s, x + (a == 0), y + (a == 1), z + (a == 2), wp);
      flag_axis(wm.rho, w0.rho, wp.rho, nr, dr);
      flag_axis(wm.p, w0.p, wp.p, np, dp);
    }
    const T chi = tmax(tsqrt(tdiv(nr, tmax(dr, T(1e-30)))), tsqrt(tdiv(np, tmax(dp, T(1e-30)))));
    r = tmax(r, tmax(chi, T(FLAG_GAS_MIN)));
  }
  r = wave_max(r);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) out[s] = tmax(tmax(red[0], red[1]), tmax(red[2], red[3]));
}

template <typename T, int LIM, int RS, bool ISEN, int D = 3>
struct AmrImpl {
  using G = AmrGeo<D>;
  static const Params<T>& P(const AmrLaunch& L) { return *static_cast<const Params<T>*>(L.params); }
  static const Levels<T>& LV(const AmrLaunch& L) { return *static_cast<const Levels<T>*>(L.levels); }

  static void assemble(const AmrLaunch& L, const int* list, int count, const void* u1, const void* u2, void* w1,
                       void* w2, unsigned long long* slot) {
    if (count <= 0) return;
    const int grid = D == 3 ? count * asm_chunks<D>() : count + (count + 1) / 2;
    hipLaunchKernelGGL((k_amr_assemble<T, LIM, D>), dim3(grid), dim3(256), 0, L.stream, P(L), LV(L), L.tab, list,
                       (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)w1, (T4<T>*)w2, slot, L.bins, count);
  }
  static void clock(const AmrLaunch& L, void* clk, unsigned long long* slot) {
    Params<T> P1 = P(L);
    P1.dx = T(1);   // PER_DX: the reduced rates are already per unit length
    hipLaunchKernelGGL(k_clock<T>, dim3(1), dim3(1), 0, L.stream, P1, (T*)clk, slot);
  }
  static void predict(const AmrLaunch& L, const int* list, int count, const void* w1, const void* w2, void* pa,
                      void* pb, const void* clk) {
    if (count <= 0) return;
    const int grid = D == 3 ? count * pred_chunks<D>() : count + (count + 3) / 4;
    hipLaunchKernelGGL((k_amr_predict<T, LIM, ISEN, D>), dim3(grid), dim3(256), 0, L.stream, P(L), LV(L), L.tab, list,
                       (const T4<T>*)w1, (const T4<T>*)w2, (T4<T>*)pa, (T4<T>*)pb, (const T*)clk, L.bins, count);
  }
  static void update(const AmrLaunch& L, const int* list, int count, bool match, bool comp, const void* u1,
                     const void* u2, const void* w1, const void* w2, const void* pa, const void* pb, void* o1, void* o2,
                     const void* clk, bool debug) {
    if (count <= 0) return;
#define SHX_UPD(M, C, DB)                                                                                              \
  hipLaunchKernelGGL((k_amr_update<T, LIM, RS, ISEN, M, C, DB, D>), dim3(count * UpdTile<D>::TILES),                   \
                     dim3(UpdTile<D>::NT), 0, L.stream, P(L), LV(L), L.tab, list, (const T4<T>*)u1, (const T4<T>*)u2,  \
                     (const T4<T>*)w1, (const T4<T>*)w2, (const T4<T>*)pa, (const T4<T>*)pb, (T4<T>*)o1, (T4<T>*)o2,   \
                     (const T*)clk, L.bins)
    if (debug) {
      if (!match) SHX_UPD(false, false, true);
      else if (comp) SHX_UPD(true, true, true);
      else SHX_UPD(true, false, true);
    } else {
      if (!match) SHX_UPD(false, false, false);
      else if (comp) SHX_UPD(true, true, false);
      else SHX_UPD(true, false, false);
    }
#undef SHX_UPD
  }
  static void rates(const AmrLaunch& L, const int* list, int count, const void* u1, const void* u2,
                    unsigned long long* slot) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_amr_rates<T, D>), dim3(count), dim3(256), 0, L.stream, P(L), LV(L), L.tab, list,
                       (const T4<T>*)u1, (const T4<T>*)u2, slot);
  }
  static void diag(const AmrLaunch& L, const DiagGeo& Gd, const int* list, int count, const void* u1, const void* u2,
                   double* part, double* out) {
    if (count > 0)
      hipLaunchKernelGGL((k_amr_diag<T, D>), dim3(count), dim3(256), 0, L.stream, P(L), Gd, L.tab, list, L.dx0,
                         (const T4<T>*)u1, (const T4<T>*)u2, part);
    hipLaunchKernelGGL(k_diag_final<T>, dim3(1), dim3(32 * DIAG_CH), 0, L.stream, (const double*)part, count, out);
  }
  static void shell(const AmrLaunch& L, const int* work0, int n0, const int* work1, int n1, const int* work2, int n2,
                    const void* u1, const void* u2, void* sh, const void* clk);
  static void fused(const AmrLaunch& L, const AmrQueue& Q, const void* sh, const void* u1, const void* u2, void* o1,
                    void* o2, const void* clk, unsigned long long* slot);
  static void regrid(const AmrLaunch& Lold, const int* map, int count, const void* u1, const void* u2, void* n1a,
                     void* n2a, void* n1b, void* n2b) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_amr_regrid<T, LIM, D>), dim3(count * (G::CELLS / 256)), dim3(256), 0, Lold.stream, P(Lold),
                       Lold.tab, (const int4*)map, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)n1a, (T4<T>*)n2a,
                       (T4<T>*)n1b, (T4<T>*)n2b);
  }
  static void indicator(const AmrLaunch& L, const int* list, int count, const void* u1, const void* u2, void* out) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_amr_indicator<T, LIM, D>), dim3(count), dim3(256), 0, L.stream, P(L), L.tab, list,
                       (const T4<T>*)u1, (const T4<T>*)u2, (T*)out);
  }
  static void block_rates(const AmrLaunch& L, const int* list, int count, const void* u1, const void* u2, void* out) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_amr_block_rates<T, D>), dim3(count), dim3(256), 0, L.stream, P(L), L.tab, list,
                       (const T4<T>*)u1, (const T4<T>*)u2, (T*)out);
  }
  static void bank(const AmrLaunch& L, void* clk) {
    hipLaunchKernelGGL((k_amr_bank<T>), dim3(1), dim3(1), 0, L.stream, (T*)clk);
  }
  static void clock_cycle(const AmrLaunch& L, void* clk, unsigned long long* slot, int cycle) {
    hipLaunchKernelGGL((k_amr_clock_cycle<T>), dim3(1), dim3(1), 0, L.stream, P(L), (T*)clk, slot, T(cycle));
  }
  static void commit(const AmrLaunch& L, const int* list, int count, const void* o1, const void* o2, void* u1,
                     void* u2) {
    if (count <= 0) return;
    hipLaunchKernelGGL((k_amr_commit<T, D>), dim3(count * (G::CELLS / 256)), dim3(256), 0, L.stream, list,
                       (const T4<T>*)o1, (const T4<T>*)o2, (T4<T>*)u1, (T4<T>*)u2);
  }
  static AmrOps ops() {
    AmrOps o;
    o.regrid = regrid;
    o.indicator = indicator;
    o.block_rates = block_rates;
    o.bank = bank;
    o.clock_cycle = clock_cycle;
    o.commit = commit;
    if constexpr (D == 3) {
      // the fused form (the ghost shells and the persistent march) is built for 3D blocks
      o.shell = shell;
      o.fused = fused;
    } else {
      o.shell = nullptr;
      o.fused = nullptr;
    }
    o.assemble = assemble;
    o.clock = clock;
    o.predict = predict;
    o.update = update;
    o.rates = rates;
    o.diag = diag;
    return o;
  }
};

}  // namespace shx
