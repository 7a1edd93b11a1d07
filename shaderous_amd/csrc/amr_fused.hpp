// The fused block-AMR step for gfx950: two batched kernels over device lists of leaves compute
// what the reference's assemble, predict and update passes compute (src/amr/assemble.wgsl,
// predict.wgsl, update.wgsl), with no W, PA or PB atlas of the blocks' cells in HBM:
//
//  - k_amr_shell: per leaf, the ghost texels of its 20^3 region, formed from U (same-level
//    neighbours, boundary images, prolongation from a coarser leaf, restriction of refined
//    children: amr_texel), and the half steps of its first ghost layer where those are another
//    block's cells ("at source", amr/predict.wgsl:183-221), all lanes busy; k_amr_rim, the half
//    steps on the rims of faces beside same-level leaves, once per source cell;
//  - k_amr_fused, one persistent kernel per launch: a workgroup owns one 16^3 block at a time
//    and marches along z: its cells from U, its ghosts from the shell; the half step of every
//    cell and of the ghosts beside a coarser block (masked transverse slopes), in LDS;
//  - every face once; the conservative update of each cell; the next step's rate / dx;
//  - coarse-fine faces through a flux register: a leaf writes the fluxes of its faces on a
//    coarser block (as it computed them), and the coarser leaf reads the four fine fluxes of
//    each of its faces there and takes their mean (TwoSum-compensated when asked), which is
//    what match_faces (amr/update.wgsl:207-290) recomputes from the finer block's atlas.
//  - one launch for all leaves: workgroups take pairs from a device queue (leaves with no
//    refined neighbour first, then the others level by level from the finest); a leaf with a
//    refined neighbour first waits for the finer leaves across it (a done flag per slot, set
//    with release order after their march). The queue hands out pairs in order and a leaf only
//    waits for leaves earlier in it, all of which are held by running workgroups, so the waits
//    end; each is bounded besides, and a wait that runs out sets an error flag the host reports.
//
// Thread layout: 384 threads (6 waves) own the columns of the block's 20 x 20 texel plane that
// ever hold a value: threads 0-255 the 16 x 16 cell columns, 256-319 the 64 face-ghost columns
// of ring 1, 320-383 the 64 of ring 2 (x or y = -2 or 17, the other in range). The 16 remaining
// columns (both x and y outside the cells) are edges or corners at every plane: zero, as in the
// atlas. So the cell waves never diverge on ghost work. LDS holds the plane data in raster order
// of the 20 x 20 columns, fluxes apart from sides (the role-tiled uniform kernel's layout, 73.6 KB
// per block in FP32), so a plane needs two barriers: after the half steps, after the fluxes.
// Two blocks share a 12-wave workgroup (one per CU at up to 168 VGPRs).
#pragma once
#include "amr_impl.hpp"

namespace shx {

constexpr int AF_SIDE = AMR_PAD;              // 20 columns per side
constexpr int AF_NT = AF_SIDE * AF_SIDE;      // 400 columns
constexpr int AF_THREADS = 384;             // one block's march
// blocks marched side by side per workgroup: two in FP32 (2 x 51 KB of LDS), one in FP64
template <typename T>
constexpr int af_pair() { return sizeof(T) == 4 ? 2 : 1; }
constexpr int AF_RING = 128;                  // ghost columns with a value: 64 of ring 1, 64 of ring 2

// thread -> column (x, y) in block coordinates, [-2, 18). A cell wave holds a 4-wide strip of the
// 16 rows: with the LDS rows 20 columns apart its lanes reach distinct banks (4 rows of 16 would
// meet in a third of them).
__device__ __forceinline__ void af_column(int t, int& x, int& y) {
  if (t < 256) { x = 4 * (t >> 6) + (t & 3); y = (t & 63) >> 2; return; }
  const int ring = t < 320 ? 1 : 2;
  t = (t - 256) & 63;
  const int side = t >> 4, i = t & 15;
  const int lo = -ring, hi = AMR_BS - 1 + ring;
  x = side == 0 ? lo : (side == 1 ? hi : i);
  y = side == 2 ? lo : (side == 3 ? hi : i);
}

// Assembled primitives of a texel of block `slot` (zero for edges and corners, as the atlas).
// LEAF: the caller knows `slot` is a leaf, whose cells are read directly (a parent's cells are
// the means of its children's, u_at).
template <typename T, int LIM, bool LEAF>
__device__ __forceinline__ Prim<T> texel_of(const Params<T>& P, const int* __restrict__ tab,
                                            const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, int slot,
                                            int x, int y, int z) {
  if (x >= 0 && x < AMR_BS && y >= 0 && y < AMR_BS && z >= 0 && z < AMR_BS && (LEAF || is_leaf(tab, slot))) {
    const size_t o = (size_t)slot * AMR_CELLS + cel(x, y, z);
    return prim_from_cons(P, U1[o], U2[o]);
  }
  Prim<T> w;
  if (!amr_texel<T, LIM>(P, tab, U1, U2, slot, x, y, z, w)) {
    w.rho = w.vx = w.vy = w.vz = w.p = w.s = w.cs = T(0);
  }
  return w;
}

// The half step of texel (x, y, z) of block `slot` from that block's texels formed directly
// (predict_cell on the source block, amr/predict.wgsl:90-181); P carries the block's dx
template <typename T, int LIM, bool ISEN>
__device__ __forceinline__ Half<T> half_at_source(const Params<T>& P, const int* __restrict__ tab,
                                               const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, int slot,
                                               int x, int y, int z, T dt) {
  const Prim<T> q = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x, y, z);
  const Prim<T> xm = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x - 1, y, z);
  const Prim<T> xp = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x + 1, y, z);
  const Prim<T> ym = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x, y - 1, z);
  const Prim<T> yp = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x, y + 1, z);
  const Prim<T> zm = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x, y, z - 1);
  const Prim<T> zp = texel_of<T, LIM, false>(P, tab, U1, U2, slot, x, y, z + 1);
  AxisSlope<T> sl[3];
  sl[0] = axis_slope<T, LIM, ISEN>(P, xm, q, xp, nullptr, nullptr);
  sl[1] = axis_slope<T, LIM, ISEN>(P, ym, q, yp, nullptr, nullptr);
  sl[2] = axis_slope<T, LIM, ISEN>(P, zm, q, zp, nullptr, nullptr);
  const T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy) + (zp.vz - zm.vz);
  const T jump = tmax(jump_of(xm.p, xp.p), tmax(jump_of(ym.p, yp.p), jump_of(zm.p, zp.p)));
  return hancock(P, q, sl, conv, jump, dt);
}

// Flux register: per coarse-fine face of a leaf, its 16 x 16 faces' fluxes (8 values each), in
// the face's transverse order (t2 * 16 + t1, t1 = (axis + 1) % 3, t2 = (axis + 2) % 3)
template <typename T>
__device__ __forceinline__ void reg_put(T* __restrict__ R, int reg, int idx, const Flux<T>& F) {
  T4<T>* p = reinterpret_cast<T4<T>*>(R + ((size_t)reg * 256 + idx) * 8);
  p[0] = v4<T>(F.m, F.mx, F.my, F.mz);
  p[1] = v4<T>(F.e, F.ei, F.s, F.vn);
}

// The mean of the four fine fluxes through the face of cell (x, y, z) on side k, read from the
// register of the refined neighbour's child; summed in match_faces' order (t2 outer, t1 inner)
template <typename T, bool COMP, bool COHERENT>
__device__ __forceinline__ Flux<T> reg_mean(const int* __restrict__ tab, const int* __restrict__ reg_of,
                                            const T* __restrict__ R, int np, int k, int x, int y, int z,
                                            const Flux<T>& fallback) {
  const int axis = k >> 1, upper = !(k & 1);
  const int hx = x >= AMR_BS / 2, hy = y >= AMR_BS / 2, hz = z >= AMR_BS / 2;
  const int bx = axis == 0 ? upper : hx, by = axis == 1 ? upper : hy, bz = axis == 2 ? upper : hz;
  const int ch = bt(tab, np, BT_CHILD + (bx | (by << 1) | (bz << 2)));
  const int reg = ch >= 0 ? reg_of[ch * 6 + (k ^ 1)] : -1;
  if (reg < 0) return fallback;   // not reached in a 2:1-balanced tree (checked at creation)
  // transverse fine coordinates of the first of the four faces
  const int fx = 2 * (x - (AMR_BS / 2) * hx), fy = 2 * (y - (AMR_BS / 2) * hy), fz = 2 * (z - (AMR_BS / 2) * hz);
  const int u0 = axis == 0 ? fy : (axis == 1 ? fz : fx);   // t1
  const int v0 = axis == 0 ? fz : (axis == 1 ? fx : fy);   // t2
  T sum[8], err[8];
#pragma unroll
  for (int q = 0; q < 8; q++) { sum[q] = T(0); err[q] = T(0); }
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const T* p = R + ((size_t)reg * 256 + (v0 + j) * 16 + (u0 + i)) * 8;
      T fv[8];
      if constexpr (COHERENT) {
        // device-coherent loads: written by another workgroup of this launch
#pragma unroll
        for (int q = 0; q < 8; q++) fv[q] = __hip_atomic_load(p + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const T4<T> a = reinterpret_cast<const T4<T>*>(p)[0], b = reinterpret_cast<const T4<T>*>(p)[1];
        fv[0] = a.x; fv[1] = a.y; fv[2] = a.z; fv[3] = a.w; fv[4] = b.x; fv[5] = b.y; fv[6] = b.z; fv[7] = b.w;
      }
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
  Flux<T> F;
  F.m = (sum[0] + err[0]) * T(0.25); F.mx = (sum[1] + err[1]) * T(0.25);
  F.my = (sum[2] + err[2]) * T(0.25); F.mz = (sum[3] + err[3]) * T(0.25);
  F.e = (sum[4] + err[4]) * T(0.25); F.ei = (sum[5] + err[5]) * T(0.25);
  F.s = (sum[6] + err[6]) * T(0.25); F.vn = (sum[7] + err[7]) * T(0.25);
  return F;
}

// Coarse-fine faces of the march, by a per-(slot, face) action: act >= 0 even: the neighbour is
// coarser, write this face's fluxes to register act / 2; odd: the neighbour is refined, take
// the face's fluxes from coarse-face buffer act / 2 (the means its march formed first); -1:
// neither.
template <typename T>
__device__ __forceinline__ Flux<T> boundary_face(int act, T* __restrict__ R, const T* __restrict__ CF, int idx,
                                                 Flux<T> F) {
  if (act < 0) return F;
  if (act & 1) {
    const T4<T>* p = reinterpret_cast<const T4<T>*>(CF + ((size_t)(act >> 1) * 256 + idx) * 8);
    const T4<T> a = p[0], b = p[1];
    F.m = a.x; F.mx = a.y; F.my = a.z; F.mz = a.w;
    F.e = b.x; F.ei = b.y; F.s = b.z; F.vn = b.w;
  } else {
    reg_put(R, act >> 1, idx, F);
  }
  return F;
}

template <typename T>
__device__ __forceinline__ void scale_inc(Inc<T>& q, T m) {
  q.rho *= m; q.vx *= m; q.vy *= m; q.vz *= m; q.p *= m; q.s *= m;
}

// ---------------- the ghost shell ----------------
// Per leaf, what its march needs from outside its own cells, formed in one batched pass with
// every lane busy: the ghost texels of its 20^3 region (the four ghost planes z = -2, -1, 16, 17
// whole, then the 128 ghost columns with a value over the 16 cell planes) and the half steps of its first
// ghost layer where they are another block's cells (64 face-ghost columns x 16 planes, then the
// two z-face ghost planes). 32 bytes per item.
constexpr int SH_TEXELS = 4 * AF_NT + AF_RING * AMR_BS;                     // 1600 + 128 * 16 = 3648
constexpr int SH_HALVES = 64 * AMR_BS + 2 * AMR_BS * AMR_BS;                 // 1024 + 512 = 1536
constexpr int SH_ITEMS = SH_TEXELS + SH_HALVES;                               // 5184
constexpr int SH_CHUNKS = (SH_ITEMS + 255) / 256;

__device__ __forceinline__ int ghost_plane_slot(int z) { return z == -2 ? 0 : (z == -1 ? 1 : (z == AMR_BS ? 2 : 3)); }

// The same-level leaf across face k of block s (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z), or -1. Its
// cells are the block's ghost texels there as they are (no image, no interpolation), so the
// march reads them directly, and the shell forms nothing for that face but the half steps of
// its rim.
__device__ __forceinline__ int same_leaf(const int* __restrict__ tab, int s, int k) {
  const int nb = bt(tab, s, BT_NB + k);
  return (nb >= 0 && is_leaf(tab, nb)) ? nb : -1;
}
// Away from the face's rim (both coordinates along the face in [1, 14]), the half step of a
// first-layer ghost of a same-level leaf uses only texels the march holds, the same ones the
// source block's own predictor uses: the march forms it in place, exactly as at its source.
__device__ __forceinline__ bool off_rim(int i, int j) {
  return i >= 1 && i <= AMR_BS - 2 && j >= 1 && j <= AMR_BS - 2;
}

// Work of the shell, one wave per item (shell index, slot, face k, chunk), listed on the host from
// the static mesh: PART 0, the ghost texels of face k (two layers, 512 texels: chunks 0-7) where
// its neighbour is not a same-level leaf; PART 1, the first layer's half steps (chunks 0-3)
// beside a boundary image. Faces on a coarser block have no PART 1 work (the march predicts
// those ghosts in place); beside a same-level leaf only the face's rim needs its half steps,
// which k_amr_rim forms from the cells of the leaf across.
template <int PART>
constexpr int sh_part_waves() { return PART == 0 ? 6 : 3; }

template <typename T, int LIM, bool ISEN, int PART>
__global__ __launch_bounds__(256, sh_part_waves<PART>()) void k_amr_shell(
    Params<T> P, Levels<T> LV, const int* __restrict__ tab, const int4* __restrict__ work, int nwork,
    const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, T4<T>* __restrict__ SH, const T* __restrict__ clk) {
  const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wv >= nwork) return;
  const int4 wk = work[wv];
  const int li = wk.x, s = wk.y, k = wk.z, code = wk.w;
  const int lane = threadIdx.x & 63;
  int x, y, z, item;
  if constexpr (PART == 0) {
    const int e = code * 64 + lane, hi = e >> 8, rem = e & 255;
    if (k < 4) {
      // ring hi + 1 column k * 16 + i, plane rem / 16
      const int col = hi * 64 + k * 16 + (rem & 15);
      z = rem >> 4;
      af_column(256 + col, x, y);
      item = 4 * AF_NT + z * AF_RING + col;
    } else {
      // ghost plane (k == 4: z = -2, -1; k == 5: z = 16, 17), cell column rem
      const int pl = (k == 4 ? 0 : 2) + hi;
      x = rem & 15;
      y = rem >> 4;
      z = pl == 0 ? -2 : (pl == 1 ? -1 : (pl == 2 ? AMR_BS : AMR_BS + 1));
      item = pl * AF_NT + (y + 2) * AF_SIDE + (x + 2);
    }
    const Prim<T> w = texel_of<T, LIM, true>(P, tab, U1, U2, s, x, y, z);
    T4<T>* out = SH + ((size_t)li * SH_ITEMS + item) * 2;
    out[0] = v4<T>(w.rho, w.vx, w.vy, w.vz);
    out[1] = v4<T>(w.p, w.s, w.cs, T(0));
  } else {
    const int e = code * 64 + lane, i = e & 15, j = e >> 4;
    int h;
    if (k < 4) {
      const int col = k * 16 + i;
      af_column(256 + col, x, y);
      z = j;
      h = col * AMR_BS + z;
    } else {
      x = i;
      y = j;
      z = k == 4 ? -1 : AMR_BS;
      h = 64 * AMR_BS + (k == 4 ? 0 : AMR_BS * AMR_BS) + y * AMR_BS + x;
    }
    const GSrc g = ghost_src(tab, s, x, y, z);
    if (g.kind != 0) return;   // beside a coarser block: predicted in place by the march
    Params<T> Pl = P;
    Pl.dx = level_dx(P, bt(tab, g.slot, BT_LEVEL));
    Half<T> hc = half_at_source<T, LIM, ISEN>(Pl, tab, U1, U2, g.slot, g.x, g.y, g.z, clk[0]);
    hc.vx = bc_vel(hc.vx, g.fx, g.sx);
    hc.vy = bc_vel(hc.vy, g.fy, g.sy);
    hc.vz = bc_vel(hc.vz, g.fz, g.sz);
    T4<T>* out = SH + ((size_t)li * SH_ITEMS + SH_TEXELS + h) * 2;
    out[0] = v4<T>(hc.rho, hc.vx, hc.vy, hc.vz);
    out[1] = v4<T>(hc.p, hc.s, hc.shock, hc.fo);
  }
}

// ---------------- the march ----------------
// Before each march: the queue starts at pair 0 and the step's epoch (the value done flags are
// set to) moves on.
template <typename T>
__global__ void k_amr_queue_start(int* qctl) {
  qctl[AMR_Q_NEXT] = 0;
  qctl[AMR_Q_EPOCH] = qctl[AMR_Q_EPOCH] + 1;
}

// Launch bound: twelve waves per workgroup (two blocks), one workgroup per CU in FP32 (three
// waves per SIMD: up to 168 VGPRs); six waves (one block) in FP64
template <typename T, int LIM, int RS, bool ISEN>
__global__ __launch_bounds__((sizeof(T) == 4 ? 2 : 1) * AF_THREADS, (sizeof(T) == 4 ? 3 : 1)) void k_amr_fused(
    Params<T> P, const int* __restrict__ tab, const int* __restrict__ list, int count,
    const int* __restrict__ shell_of, const T4<T>* __restrict__ SH, const int* __restrict__ act,
    T* __restrict__ R, T* __restrict__ CF, const int* __restrict__ reg_of, int comp, const int* __restrict__ dep_off,
    const int* __restrict__ dep, int* __restrict__ qctl, int* __restrict__ done, const T4<T>* __restrict__ U1,
    const T4<T>* __restrict__ U2, T4<T>* __restrict__ O1, T4<T>* __restrict__ O2, const T* __restrict__ clk,
    unsigned long long* next_slot) {
  constexpr int NT = AF_NT, BX = AF_SIDE, AF_PAIR = af_pair<T>();
  using LY = RoleLayout<NT>;
  __shared__ T LL[AF_PAIR * LY::SIZE];
  __shared__ T red[AF_PAIR * AF_THREADS / 64];
  __shared__ int q_pair;
  const int epoch = qctl[AMR_Q_EPOCH];
  // which of the workgroup's two blocks this thread marches, and its thread index there
  const int half = threadIdx.x >= AF_THREADS;
  const int tid = threadIdx.x - AF_THREADS * half;
  T* L = LL + half * LY::SIZE;
  int x, y;
  af_column(tid, x, y);
  const int t = (y + 2) * BX + (x + 2);                  // raster index of the column in LDS
  // the columns no thread owns are edges at every plane: zero in both plane slots, for good
  for (int i = tid; i < 12 * NT; i += AF_THREADS) L[i] = T(0);
  // roles by wave (uniform branches): waves 0-3 the cell columns, wave 4 the face ghosts of
  // ring 1, wave 5 ring 2 (loads only)
  const int wave = __builtin_amdgcn_readfirstlane(tid) >> 6;
  const bool cells0 = wave < 4, ghosts0 = wave == 4;
  // a face-ghost lane's side: 0 -x, 1 +x, 2 -y, 3 +y
  const int side = (tid >> 4) & 3;
  const T dt = clk[0];
  const T dx0 = P.dx;
  T rate_max = T(0);

  for (;;) {
    // the next pair from the queue (every thread has passed a barrier since it last read q_pair)
    if (threadIdx.x == 0) q_pair = atomicAdd(&qctl[AMR_Q_NEXT], 1);
    __syncthreads();
    const int wp = q_pair;
    __syncthreads();   // every thread has read q_pair before thread 0 takes the next ticket
    if (AF_PAIR * wp >= count) break;
    // -1 pads a queue part of odd length. With pairs it is always a pair's second entry (parts
    // start at even positions): that half takes part in the barriers only. A one-block
    // workgroup skips it.
    if (list[AF_PAIR * wp] < 0) continue;
    const int wi = AF_PAIR * wp + half;
    const bool act_blk = wi < count && list[wi] >= 0;
    const int s = list[act_blk ? wi : AF_PAIR * wp];
    // a leaf with a refined neighbour: wait for the finer leaves across it (one lane each), then
    // form the means of their fluxes on its coarse-fine faces
    const int d0 = act_blk ? dep_off[s] : 0, nd = act_blk ? dep_off[s + 1] - d0 : 0;
    // The flags are read with agent-scope acquire loads, pairing with the writers' release
    // fence: what a finer leaf wrote before its flag (its register fluxes) is visible to the
    // reads after the barrier below, which are device-coherent loads as well
    if (tid < nd) {
      int* flag = done + dep[d0 + tid];
      int spins = 0;
      while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > AMR_WAIT_SPINS) {
          __hip_atomic_store(&qctl[AMR_Q_ERROR], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __syncthreads();
    if (nd > 0) {
      for (int k = 0; k < 6; k++) {
        const int a = act[s * 6 + k];
        if (a < 0 || !(a & 1) || tid >= 256) continue;
        const int axis = k >> 1, t1 = tid & 15, t2 = tid >> 4;
        const int n = (k & 1) ? AMR_BS - 1 : 0;
        const int fx = axis == 0 ? n : (axis == 1 ? t2 : t1);
        const int fy = axis == 1 ? n : (axis == 2 ? t2 : t1);
        const int fz = axis == 2 ? n : (axis == 0 ? t2 : t1);
        const int np = bt(tab, s, BT_NB + k);
        const Flux<T> F = comp ? reg_mean<T, true, true>(tab, reg_of, R, np, k, fx, fy, fz, zero_flux<T>())
                               : reg_mean<T, false, true>(tab, reg_of, R, np, k, fx, fy, fz, zero_flux<T>());
        T4<T>* q = reinterpret_cast<T4<T>*>(CF + ((size_t)(a >> 1) * 256 + tid) * 8);
        q[0] = v4<T>(F.m, F.mx, F.my, F.mz);
        q[1] = v4<T>(F.e, F.ei, F.s, F.vn);
      }
    }
    const T4<T>* sh = SH + (size_t)shell_of[s] * SH_ITEMS * 2;
    // this block's cell size (P is the kernel's own copy: its dx is the block's from here on)
    const T dx = dx0 * (T(1) / T(1 << bt(tab, s, BT_LEVEL)));
    P.dx = dx;
    const T dtdx = dt / dx;
    const bool cells = cells0 && act_blk, ghosts = ghosts0 && act_blk;
    // beside a coarser block (-1): its ghosts are predicted in place; elsewhere they are another
    // block's cells, whose half steps come from the shell
    const bool gk = ghosts && bt(tab, s, BT_NB + side) == -1;
    const bool kz_lo = bt(tab, s, BT_NB + 4) == -1, kz_hi = bt(tab, s, BT_NB + 5) == -1;
    // same-level leaves: across this ring lane's side, and across the two z faces
    const int nbs = (!cells0 && act_blk) ? same_leaf(tab, s, side) : -1;
    const int nbz_lo = same_leaf(tab, s, 4), nbz_hi = same_leaf(tab, s, 5);
    __syncthreads();   // the previous block's last LDS reads are done

    // texel (x, y, z) of this block: its own cells and a same-level leaf's from U, edges zero,
    // other ghosts from the shell
    auto texel = [&](int z) -> Prim<T> {
      Prim<T> w;
      w.rho = w.vx = w.vy = w.vz = w.p = w.s = w.cs = T(0);
      const bool oz = z < 0 || z >= AMR_BS;
      if (!act_blk || (!cells0 && oz)) return w;
      int src = -1;
      if (cells0) {
        if (!oz) src = s;
        else if ((z < 0 ? nbz_lo : nbz_hi) >= 0) src = z < 0 ? nbz_lo : nbz_hi;
      } else {
        src = nbs;
      }
      if (src >= 0) {
        const int xs = x < 0 ? x + AMR_BS : (x >= AMR_BS ? x - AMR_BS : x);
        const int ys = y < 0 ? y + AMR_BS : (y >= AMR_BS ? y - AMR_BS : y);
        const int zs = z < 0 ? z + AMR_BS : (z >= AMR_BS ? z - AMR_BS : z);
        const size_t o = (size_t)src * AMR_CELLS + cel(xs, ys, zs);
        return prim_from_cons(P, U1[o], U2[o]);
      }
      const int g = oz ? ghost_plane_slot(z) * AF_NT + t : 4 * AF_NT + z * AF_RING + (tid - 256);
      const T4<T> a = sh[2 * g], b = sh[2 * g + 1];
      w.rho = a.x; w.vx = a.y; w.vy = a.z; w.vz = a.w;
      w.p = b.x; w.s = b.y; w.cs = b.z;
      return w;
    };
    lds_put_prim6<T, NT>(L + LY::Q0, t, texel(-2));
    T cs_c;   // this column's sound speed in plane c
    {
      const Prim<T> q = texel(-1);
      lds_put_prim6<T, NT>(L + LY::Q1, t, q);
      cs_c = q.cs;
    }
    Prim<T> nxt = texel(0);
    __syncthreads();

    SideData<T> zs_prev;
    Flux<T> fz_lo = zero_flux<T>();
    Flux<T> axy_prev = zero_flux<T>();

    for (int c = -1; c <= AMR_BS; c++) {
      const Prim<T> pp = nxt;             // plane c + 1
      if (c + 2 <= AMR_BS + 1) nxt = texel(c + 2);
      const bool odd = ((c + 1) & 1) != 0;
      const T* Lq = L + (odd ? LY::Q0 : LY::Q1);
      T* Lm = L + (odd ? LY::Q1 : LY::Q0);

      // 2. the half step of texel (x, y, c) and its sides: every cell; a ghost plane's texels
      //    (c = -1, 16) in the cell waves; the face ghosts of the cell planes in the ghost wave
      const bool oz = c < 0 || c >= AMR_BS;
      SideData<T> sxl, syl, szl, szu;
      if (cells || (ghosts && !oz)) {
        Prim<T> pc = lds_prim<T, NT>(Lq, t);
        pc.cs = cs_c;
        const Prim<T> pm = lds_prim<T, NT>(Lm, t);
        const Prim<T> xm = lds_prim<T, NT>(Lq, t - 1), xp = lds_prim<T, NT>(Lq, t + 1);
        const Prim<T> ym = lds_prim<T, NT>(Lq, t - BX), yp = lds_prim<T, NT>(Lq, t + BX);
        Inc<T> ixu, ixd, iyu, iyd, izu, izd;
        AxisSlope<T> sl[3];
        sl[0] = axis_slope<T, LIM, ISEN>(P, xm, pc, xp, &ixu, &ixd);
        sl[1] = axis_slope<T, LIM, ISEN>(P, ym, pc, yp, &iyu, &iyd);
        sl[2] = axis_slope<T, LIM, ISEN>(P, pm, pc, pp, &izu, &izd);
        const bool ghost = ghosts || oz;
        const bool in_place = ghosts ? gk : (oz && (c < 0 ? kz_lo : kz_hi));
        // a same-level leaf's cell away from the face's rim: as at its source, in place
        const bool exact = ghosts ? (nbs >= 0 && off_rim(side < 2 ? y : x, c))
                                  : (oz && (c < 0 ? nbz_lo : nbz_hi) >= 0 && off_rim(x, y));
        Half<T> hc;
        if (!ghost || in_place || exact) {
          if (ghost && !exact) {
            // beside a coarser block: predicted in place, the slopes along the ghost's rim axes
            // dropped where their neighbour is an unassembled edge texel (amr/predict.wgsl:183-221)
            const bool ox = x < 0 || x >= AMR_BS, oy = y < 0 || y >= AMR_BS;
            scale_inc(sl[0].s, (!ox && (x == 0 || x == AMR_BS - 1)) ? T(0) : T(1));
            scale_inc(sl[1].s, (!oy && (y == 0 || y == AMR_BS - 1)) ? T(0) : T(1));
            scale_inc(sl[2].s, (!oz && (c == 0 || c == AMR_BS - 1)) ? T(0) : T(1));
          }
          const T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy) + (pp.vz - pm.vz);
          const T jump = tmax(jump_of(xm.p, xp.p), tmax(jump_of(ym.p, yp.p), jump_of(pm.p, pp.p)));
          hc = hancock(P, pc, sl, conv, jump, dt);
        } else {
          // another block's cell (same level, refined, or this block's boundary image): its half
          // step was formed in the shell pass
          const int h = oz ? 64 * AMR_BS + (c < 0 ? 0 : AMR_BS * AMR_BS) + y * AMR_BS + x : (tid - 256) * AMR_BS + c;
          const T4<T> a = sh[2 * (SH_TEXELS + h)], b = sh[2 * (SH_TEXELS + h) + 1];
          hc.rho = a.x; hc.vx = a.y; hc.vy = a.z; hc.vz = a.w;
          hc.p = b.x; hc.s = b.y; hc.shock = b.z; hc.fo = b.w;
        }
        sxl = side_of(hc, ixd, T(-1));
        syl = side_of(hc, iyd, T(-1));
        szl = side_of(hc, izd, T(-1));
        szu = side_of(hc, izu, T(1));
        lds_put_side<T, NT>(L, LY::SX, t, side_of(hc, ixu, T(1)));
        lds_put_side<T, NT>(L, LY::SY, t, side_of(hc, iyu, T(1)));
      }
      // plane c + 1 into the slot plane c - 1 leaves
      lds_put_prim6<T, NT>(Lm, t, pp);
      cs_c = pp.cs;
      __syncthreads();

      // 3. the x and y faces below this column (cells; the +x and +y ghosts' lower faces), and
      //    the z face below plane c
      const bool live = c >= 0 && c < AMR_BS;
      Flux<T> Fx = zero_flux<T>(), Fy = zero_flux<T>(), fz = zero_flux<T>();
      if (live && (cells || (ghosts && side == 1))) {
        Fx = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SX, t - 1), sxl, 0);
        if (x == 0) Fx = boundary_face<T>(act[s * 6 + 0], R, CF, c * 16 + y, Fx);
        else if (x == AMR_BS) Fx = boundary_face<T>(act[s * 6 + 1], R, CF, c * 16 + y, Fx);
      }
      if (live && (cells || (ghosts && side == 3))) {
        Fy = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SY, t - BX), syl, 1);
        if (y == 0) Fy = boundary_face<T>(act[s * 6 + 2], R, CF, x * 16 + c, Fy);
        else if (y == AMR_BS) Fy = boundary_face<T>(act[s * 6 + 3], R, CF, x * 16 + c, Fy);
      }
      if (cells && c >= 0) {
        fz = face_from_sides<T, RS>(P, zs_prev, szl, 2);
        if (c == 0) fz = boundary_face<T>(act[s * 6 + 4], R, CF, y * 16 + x, fz);
        else if (c == AMR_BS) fz = boundary_face<T>(act[s * 6 + 5], R, CF, y * 16 + x, fz);
      }
      // 4. the x/y fluxes into LDS
      if (live && (cells || (ghosts && side == 1))) lds_put_flux<T, NT>(L, LY::FX, t, Fx);
      if (live && (cells || (ghosts && side == 3))) lds_put_flux<T, NT>(L, LY::FY, t, Fy);
      __syncthreads();

      // 5. this plane's x/y net inflow, and the update of plane c - 1
      if (cells) {
        Flux<T> axy = zero_flux<T>();
        if (live) {
          const Flux<T> xh = lds_get_flux<T, NT>(L, LY::FX, t + 1);
          const Flux<T> yh = lds_get_flux<T, NT>(L, LY::FY, t + BX);
          axy.m = (Fx.m - xh.m) + (Fy.m - yh.m);
          axy.mx = (Fx.mx - xh.mx) + (Fy.mx - yh.mx);
          axy.my = (Fx.my - xh.my) + (Fy.my - yh.my);
          axy.mz = (Fx.mz - xh.mz) + (Fy.mz - yh.mz);
          axy.e = (Fx.e - xh.e) + (Fy.e - yh.e);
          axy.ei = (Fx.ei - xh.ei) + (Fy.ei - yh.ei);
          axy.s = (Fx.s - xh.s) + (Fy.s - yh.s);
          axy.vn = (Fx.vn - xh.vn) + (Fy.vn - yh.vn);
        }
        if (c - 1 >= 0 && c - 1 < AMR_BS) {
          const size_t o = (size_t)s * AMR_CELLS + cel(x, y, c - 1);
          Flux<T> net;
          net.m = axy_prev.m + (fz_lo.m - fz.m);
          net.mx = axy_prev.mx + (fz_lo.mx - fz.mx);
          net.my = axy_prev.my + (fz_lo.my - fz.my);
          net.mz = axy_prev.mz + (fz_lo.mz - fz.mz);
          net.e = axy_prev.e + (fz_lo.e - fz.e);
          net.ei = axy_prev.ei + (fz_lo.ei - fz.ei);
          net.s = axy_prev.s + (fz_lo.s - fz.s);
          net.vn = axy_prev.vn + (fz_lo.vn - fz.vn);
          T4<T> o1, o2;
          advance(P, U1[o], U2[o], net, zs_prev.hp, dtdx, o1, o2);
          O1[o] = o1;
          O2[o] = o2;
          // the next step's rate / dx (assemble's leaf-cell rate, amr/assemble.wgsl:77-154)
          rate_max = tmax(rate_max, rate_of(P, prim_from_cons(P, o1, o2)) / dx);
        }
        zs_prev = szu;
        fz_lo = fz;
        axy_prev = axy;
      }
    }
    // a leaf on a coarser block: its register fluxes are written (the barrier waits for every
    // wave's stores); one release fence writes them back to device-coherent memory, then its
    // done flag
    __syncthreads();
    bool writes_reg = false;
    for (int k = 0; k < 6; k++) {
      const int a = act[s * 6 + k];
      writes_reg = writes_reg || (a >= 0 && !(a & 1));
    }
    if (tid == 0 && act_blk && writes_reg) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(done + s, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  // workgroup maximum of the next step's rate, one atomicMax
  T v = wave_max(rate_max);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T m = T(0);
    for (int w = 0; w < AF_PAIR * AF_THREADS / 64; w++) m = tmax(m, red[w]);
    max_slot(next_slot, (unsigned long long)order_bits(m));
  }
}

// The half steps of the face-rim ghosts beside same-level leaves: a rim ghost's source is a cell
// on an edge of the leaf across the face, whose transverse stencil reaches past the face (the
// march holds only the face itself). One lane per source cell, from records listed on the host
// (source slot, cell x | y << 4 | z << 8, then up to three (shell index, half-step item)
// destinations, -1 for none): an edge cell feeds the shells of the leaves across both its faces,
// a corner cell three, so each is predicted once, and the lanes of a wave take the edge cells of
// one block. The value is the shell's own for that ghost (half_at_source at the source; a
// same-level face maps no velocity).
template <typename T, int LIM, bool ISEN>
__global__ __launch_bounds__(256) void k_amr_rim(Params<T> P, const int* __restrict__ tab,
                                                 const int4* __restrict__ work, int nwork,
                                                 const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2,
                                                 T4<T>* __restrict__ SH, const T* __restrict__ clk) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= nwork) return;
  const int4 a = work[2 * w], b = work[2 * w + 1];
  const int x = a.y & 15, y = (a.y >> 4) & 15, z = (a.y >> 8) & 15;
  Params<T> Pl = P;
  Pl.dx = level_dx(P, bt(tab, a.x, BT_LEVEL));
  const Half<T> hc = half_at_source<T, LIM, ISEN>(Pl, tab, U1, U2, a.x, x, y, z, clk[0]);
  const T4<T> v0 = v4<T>(hc.rho, hc.vx, hc.vy, hc.vz), v1 = v4<T>(hc.p, hc.s, hc.shock, hc.fo);
  const int li[3] = {a.z, b.x, b.z}, h[3] = {a.w, b.y, b.w};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (li[d] < 0) continue;
    T4<T>* out = SH + ((size_t)li[d] * SH_ITEMS + SH_TEXELS + h[d]) * 2;
    out[0] = v0;
    out[1] = v1;
  }
}

// (3D blocks only: the ops of a 2D grid leave the fused form out)
template <typename T, int LIM, int RS, bool ISEN, int D>
void AmrImpl<T, LIM, RS, ISEN, D>::shell(const AmrLaunch& L, const int* work0, int n0, const int* work1, int n1,
                                         const int* work2, int n2, const void* u1, const void* u2, void* sh,
                                         const void* clk) {
  static_assert(D == 3, "the fused form steps 3D blocks");
  if (n2 > 0)
    hipLaunchKernelGGL((k_amr_rim<T, LIM, ISEN>), dim3((n2 + 255) / 256), dim3(256), 0, L.stream, P(L), L.tab,
                       (const int4*)work2, n2, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)sh, (const T*)clk);
  if (n0 > 0)
    hipLaunchKernelGGL((k_amr_shell<T, LIM, ISEN, 0>), dim3((n0 + 3) / 4), dim3(256), 0, L.stream, P(L), LV(L), L.tab,
                       (const int4*)work0, n0, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)sh, (const T*)clk);
  if (n1 > 0)
    hipLaunchKernelGGL((k_amr_shell<T, LIM, ISEN, 1>), dim3((n1 + 3) / 4), dim3(256), 0, L.stream, P(L), LV(L), L.tab,
                       (const int4*)work1, n1, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)sh, (const T*)clk);
}

template <typename T, int LIM, int RS, bool ISEN, int D>
void AmrImpl<T, LIM, RS, ISEN, D>::fused(const AmrLaunch& L, const AmrQueue& Q, const void* sh, const void* u1,
                                         const void* u2, void* o1, void* o2, const void* clk,
                                         unsigned long long* slot) {
  static_assert(D == 3, "the fused form steps 3D blocks");
  if (Q.count <= 0) return;
  constexpr int np = af_pair<T>();
  const int pairs = (Q.count + np - 1) / np;
  const int grid = Q.grid < pairs ? Q.grid : pairs;
  hipLaunchKernelGGL((k_amr_queue_start<T>), dim3(1), dim3(1), 0, L.stream, Q.qctl);
  hipLaunchKernelGGL((k_amr_fused<T, LIM, RS, ISEN>), dim3(grid), dim3(np * AF_THREADS), 0, L.stream, P(L), L.tab,
                     Q.list, Q.count, Q.shell_of, (const T4<T>*)sh, Q.act, (T*)Q.reg, (T*)Q.cf, Q.reg_of, Q.comp,
                     Q.dep_off, Q.dep, Q.qctl, Q.done, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)o1, (T4<T>*)o2,
                     (const T*)clk, slot);
}

}  // namespace shx
