// Uniform-grid hydro step kernels for gfx950.
//
// PLAIN path, one kernel per reference pass (src/sim.ts encodeSteps :3187-3328):
//   k_rates   hydro/rates.wgsl:33-95 + the MAX reduction (gpu/reduce.wgsl): wave and block
//             reductions, one atomicMax per block on the ordered bit pattern
//   k_clock   hydro/clock.wgsl:132-242 on one thread; clears the reduction slot
//   k_predict hydro/predict.wgsl:32-119: the Hancock half step per cell -> PA, PB
//   k_update  hydro/update.wgsl:189-405 with the face tile (hydro/tile.wgsl:166-213): every
//             face of the block's box computed once into LDS, then each cell's update
//
// FUSED path, one kernel per step (k_fused): a block owns a column tile of (BX - 4) x (BY - 4)
// cells plus a two-cell halo ring and marches along z through a chunk of planes. Each plane's
// primitives are converted once, staged in LDS for the x/y neighbours and kept in registers for
// the z neighbours; each cell's half step and edge increments are computed once; each face is
// solved once; the conservative update of the plane behind is written, and the next step's
// signal rate is folded into a per-block maximum in the same pass. HBM traffic is one read and
// one write of the state per step plus the halo re-reads the caches do not absorb.
//
// Both paths evaluate the same device functions (hydro.hpp) in the same order on the same
// inputs, so their results agree bit for bit.
#pragma once
#include "uniform.hpp"

namespace shx {

// ---------------- reductions ----------------
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = tmax(v, __shfl_xor(v, o, 64));
  return v;
}

// atomicMax on a rate slot, which only grows during a launch (it is cleared before the launch
// that fills it, and a launch starts with its caches invalidated): a value not above the slot's
// current one, read from L2, changes nothing and skips the atomic. With thousands of workgroups
// or waves (the block passes, the row march) the atomics on one address queue behind each other.
__device__ __forceinline__ void max_slot(unsigned long long* slot, unsigned long long bits) {
  if (bits > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(slot, bits);
}

// Max over the block, then one atomicMax on the slot (ordered bit patterns of non-negative T).
template <typename T>
__device__ __forceinline__ void block_max_to(T v, unsigned long long* slot) {
  __shared__ T red[16];
  v = wave_max(v);
  int lane = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  int nthreads = blockDim.x * blockDim.y * blockDim.z;
  if ((lane & 63) == 0) red[lane >> 6] = v;
  __syncthreads();
  if (lane == 0) {
    T m = T(0);
    for (int w = 0; w < (nthreads + 63) / 64; w++) m = tmax(m, red[w]);
    max_slot(slot, (unsigned long long)order_bits(m));
  }
}

template <typename T>
__device__ __forceinline__ T from_bits(unsigned long long b);
template <> __device__ __forceinline__ float from_bits<float>(unsigned long long b) { return __uint_as_float((uint32_t)b); }
template <> __device__ __forceinline__ double from_bits<double>(unsigned long long b) { return __longlong_as_double((long long)b); }

// ---------------- plain path ----------------
template <typename T>
__global__ __launch_bounds__(256) void k_rates(Params<T> P, const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2,
                                               int ncell, unsigned long long* slot) {
  T m = T(0);
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < ncell; o += gridDim.x * blockDim.x) {
    T4<T> u2 = U2[o];
    Prim<T> w = prim_from_cons(P, U1[o], u2);
    T r = rate_of(P, w);
    if (u2.w <= T(0.5)) m = tmax(m, r);
  }
  block_max_to(m, slot);
}

template <typename T>
__global__ void k_clock(Params<T> P, T* clk, unsigned long long* slot) {
  T rate = from_bits<T>(*slot);
  clock_pass(P, clk, rate);
  *slot = 0ull;
}

// ---------------- the clock inside a fused step ----------------
// A fused step takes the clock pass itself instead of a clock launch before it (one launch per
// step). Each instance keeps two clocks (16 values each), one per parity of the state, three rate
// slots and two epochs (steps taken since the rates were last reset): the step from parity p
// reads clock p and the slot of its epoch e (mod 3), the largest rate of the state it reads;
// every workgroup forms the step's dt from them, as the clock pass (hydro/clock.wgsl:132-242)
// does, on a copy; the rates of the state it writes go to slot e + 1. The first workgroup also
// writes the clock pass into clock 1 - p (what the clock launch would have left in the clock:
// the dt being taken, the time it starts from), epoch e + 1, and clears slot e + 2, last read by
// the step before and next written by the step after. No word is both read and written within
// one step.
template <typename T>
struct StepClock {
  T* clk;                        // per instance: clocks of parity 0 and 1, 16 values each
  unsigned long long* slot;      // per instance: 3 rate slots
  unsigned int* epoch;           // per instance: epochs of parity 0 and 1
  int p;                         // the parity of the state the step reads
  __device__ __forceinline__ const T* clock_in(int i) const { return clk + 32 * i + 16 * p; }
  __device__ __forceinline__ unsigned int epoch_in(int i) const { return epoch[2 * i + p]; }
  __device__ __forceinline__ unsigned long long* slot_at(int i, unsigned int e) const { return slot + 3 * i + e % 3u; }
  // where the rates of the state the step writes go
  __device__ __forceinline__ unsigned long long* next(int i) const { return slot_at(i, epoch_in(i) + 1u); }
};

template <typename T>
StepClock<T> step_clock(const ClockRefs& r) {
  return StepClock<T>{static_cast<T*>(r.clk), r.slot, r.epoch, r.p};
}

// A value every lane of the wave holds alike, as a scalar register (so that it does not take a
// vector register for the rest of the kernel).
__device__ __forceinline__ float wave_uniform(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// The clock pass of instance i on a copy of its clock: c[0] is the dt of the step.
template <typename T>
__device__ __forceinline__ void step_clock_pass(const Params<T>& P, const StepClock<T>& C, int i, T c[11]) {
  const T* in = C.clock_in(i);
  for (int k = 0; k < 11; k++) c[k] = in[k];
  clock_pass(P, c, from_bits<T>(*C.slot_at(i, C.epoch_in(i))));
}

// The dt of the step instance i takes. i is the same for every lane of a wave.
template <typename T>
__device__ __forceinline__ T step_dt(const Params<T>& P, const StepClock<T>& C, int i) {
  T c[11];
  step_clock_pass(P, C, i, c);
  return wave_uniform(c[0]);
}

// By one thread of the step, for instance i: the clock the next step reads, its epoch, and the
// rate slot the next step writes cleared.
template <typename T>
__device__ __forceinline__ void step_clock_advance(const Params<T>& P, const StepClock<T>& C, int i) {
  T c[11];
  step_clock_pass(P, C, i, c);
  T* out = C.clk + 32 * i + 16 * (1 - C.p);
  for (int k = 0; k < 11; k++) out[k] = c[k];
  const unsigned int e = C.epoch_in(i);
  C.epoch[2 * i + 1 - C.p] = e + 1u;
  *C.slot_at(i, e + 2u) = 0ull;
}

template <typename T>
__device__ __forceinline__ size_t cell_index(const Params<T>& P, int i, int j, int k) {
  return ((size_t)k * P.n[1] + j) * P.n[0] + i;
}

// load_q (hydro/recon.wgsl:37-52): primitives of a possibly out-of-range cell
template <typename T>
__device__ __forceinline__ Prim<T> load_q(const Params<T>& P, const T4<T>* U1, const T4<T>* U2, int i, int j, int k) {
  AxisMap mx = map_axis(i, P.n[0], P.bc_lo[0], P.bc_hi[0]);
  AxisMap my = map_axis(j, P.n[1], P.bc_lo[1], P.bc_hi[1]);
  AxisMap mz = map_axis(k, P.n[2], P.bc_lo[2], P.bc_hi[2]);
  size_t o = cell_index(P, mx.i, my.i, mz.i);
  Prim<T> q = prim_from_cons(P, U1[o], U2[o]);
  q.vx = bc_vel(q.vx, mx.flip, mx.side);
  q.vy = bc_vel(q.vy, my.flip, my.side);
  q.vz = bc_vel(q.vz, mz.flip, mz.side);
  return q;
}

// load_pred (hydro/update.wgsl:47-63): half-step state of a possibly out-of-range cell
template <typename T>
__device__ __forceinline__ Half<T> load_half(const Params<T>& P, const T4<T>* PA, const T4<T>* PB, int i, int j, int k) {
  AxisMap mx = map_axis(i, P.n[0], P.bc_lo[0], P.bc_hi[0]);
  AxisMap my = map_axis(j, P.n[1], P.bc_lo[1], P.bc_hi[1]);
  AxisMap mz = map_axis(k, P.n[2], P.bc_lo[2], P.bc_hi[2]);
  size_t o = cell_index(P, mx.i, my.i, mz.i);
  T4<T> a = PA[o], b = PB[o];
  Half<T> h;
  h.rho = a.x;
  h.vx = bc_vel(a.y, mx.flip, mx.side);
  h.vy = bc_vel(a.z, my.flip, my.side);
  h.vz = bc_vel(a.w, mz.flip, mz.side);
  h.p = b.x; h.s = b.y; h.shock = b.z; h.fo = b.w;
  return h;
}

// The half step of cell (i, j, k) (predict.wgsl:32-119): PA = (rho, v), PB = (p, s, shock, first order)
template <typename T, int LIM, bool ISEN>
__device__ __forceinline__ void predict_cell(const Params<T>& P, const T4<T>* __restrict__ U1,
                                             const T4<T>* __restrict__ U2, int i, int j, int k, T dt, T4<T>& a,
                                             T4<T>& b) {
  Prim<T> q = load_q(P, U1, U2, i, j, k);
  Prim<T> xm = load_q(P, U1, U2, i - 1, j, k), xp = load_q(P, U1, U2, i + 1, j, k);
  Prim<T> ym = load_q(P, U1, U2, i, j - 1, k), yp = load_q(P, U1, U2, i, j + 1, k);
  if (P.flat) {
    const Half<T> h = half_flat<T, LIM, ISEN>(P, q, xm, xp, ym, yp, dt, nullptr, nullptr, nullptr, nullptr);
    a.x = h.rho; a.y = h.vx; a.z = h.vy; a.w = h.vz;
    b.x = h.p; b.y = h.s; b.z = h.shock; b.w = h.fo;
    return;
  }
  Prim<T> zm = load_q(P, U1, U2, i, j, k - 1), zp = load_q(P, U1, U2, i, j, k + 1);
  AxisSlope<T> sl[3];
  sl[0] = axis_slope<T, LIM, ISEN>(P, xm, q, xp, nullptr, nullptr);
  sl[1] = axis_slope<T, LIM, ISEN>(P, ym, q, yp, nullptr, nullptr);
  sl[2] = axis_slope<T, LIM, ISEN>(P, zm, q, zp, nullptr, nullptr);
  T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy) + (zp.vz - zm.vz);
  T jump = tmax(jump_of(xm.p, xp.p), tmax(jump_of(ym.p, yp.p), jump_of(zm.p, zp.p)));
  Half<T> h = hancock(P, q, sl, conv, jump, dt);
  a.x = h.rho; a.y = h.vx; a.z = h.vy; a.w = h.vz;
  b.x = h.p; b.y = h.s; b.z = h.shock; b.w = h.fo;
}

template <typename T, int LIM, bool ISEN>
__global__ __launch_bounds__(256) void k_predict(Params<T> P, const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2,
                                                 T4<T>* __restrict__ PA, T4<T>* __restrict__ PB, const T* clk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z * blockDim.z + threadIdx.z;
  if (i >= P.n[0] || j >= P.n[1] || k >= P.n[2]) return;
  T4<T> a, b;
  predict_cell<T, LIM, ISEN>(P, U1, U2, i, j, k, clk[0], a, b);
  size_t o = cell_index(P, i, j, k);
  PA[o] = a;
  PB[o] = b;
}

// The half steps of the first and last planes (z = 0 and nz - 1) into ZA, ZB (two planes each):
// what the fused march takes for its ghost planes beyond a non-periodic z face (load_pred,
// hydro/update.wgsl:47-63, maps a ghost to one of these cells). Grid (x, y, 2).
template <typename T, int LIM, bool ISEN>
__global__ __launch_bounds__(256) void k_zplanes(Params<T> P, const T4<T>* __restrict__ U1,
                                                 const T4<T>* __restrict__ U2, T4<T>* __restrict__ ZA,
                                                 T4<T>* __restrict__ ZB, StepClock<T> C) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n[0] || j >= P.n[1]) return;
  const int k = blockIdx.z == 0 ? 0 : P.n[2] - 1;
  T4<T> a, b;
  predict_cell<T, LIM, ISEN>(P, U1, U2, i, j, k, step_dt(P, C, 0), a, b);
  const size_t o = (size_t)blockIdx.z * P.n[0] * P.n[1] + (size_t)j * P.n[0] + i;
  ZA[o] = a;
  ZB[o] = b;
}

// flux_between (hydro/flux.wgsl:254-347): the face between cell a and a + e_axis
template <typename T, int LIM, int RS, bool ISEN>
__device__ __forceinline__ Flux<T> flux_between(const Params<T>& P, const T4<T>* U1, const T4<T>* U2, const T4<T>* PA,
                                                const T4<T>* PB, int i, int j, int k, int axis) {
  int ex = axis == 0, ey = axis == 1, ez = axis == 2;
  Prim<T> wll = load_q(P, U1, U2, i - ex, j - ey, k - ez);
  Prim<T> wl = load_q(P, U1, U2, i, j, k);
  Prim<T> wr = load_q(P, U1, U2, i + ex, j + ey, k + ez);
  Prim<T> wrr = load_q(P, U1, U2, i + 2 * ex, j + 2 * ey, k + 2 * ez);
  Half<T> hl = load_half(P, PA, PB, i, j, k);
  Half<T> hr = load_half(P, PA, PB, i + ex, j + ey, k + ez);
  Inc<T> il = edge<T, LIM, ISEN, true>(P, wll, wl, wr);
  Inc<T> ir = edge<T, LIM, ISEN, false>(P, wl, wr, wrr);
  return face_flux<T, RS>(P, hl, il, hr, ir, axis);
}

template <typename T>
__device__ __forceinline__ void put_flux(T* tf, int nf, int f, const Flux<T>& F) {
  tf[0 * nf + f] = F.m; tf[1 * nf + f] = F.mx; tf[2 * nf + f] = F.my; tf[3 * nf + f] = F.mz;
  tf[4 * nf + f] = F.e; tf[5 * nf + f] = F.ei; tf[6 * nf + f] = F.s; tf[7 * nf + f] = F.vn;
}
template <typename T>
__device__ __forceinline__ Flux<T> get_flux(const T* tf, int nf, int f) {
  Flux<T> F;
  F.m = tf[0 * nf + f]; F.mx = tf[1 * nf + f]; F.my = tf[2 * nf + f]; F.mz = tf[3 * nf + f];
  F.e = tf[4 * nf + f]; F.ei = tf[5 * nf + f]; F.s = tf[6 * nf + f]; F.vn = tf[7 * nf + f];
  return F;
}
template <typename T>
__device__ __forceinline__ Flux<T> zero_flux() {
  Flux<T> F;
  F.m = F.mx = F.my = F.mz = F.e = F.ei = F.s = F.vn = T(0);
  return F;
}

// The tiled update: block (WX, WY, WZ) cells; WZ = 1 with P.flat for 2D.
template <typename T, int LIM, int RS, bool ISEN, int WX, int WY, int WZ, bool DEBUG>
__global__ __launch_bounds__(WX* WY* WZ) void k_update(Params<T> P, const T4<T>* __restrict__ U1,
                                                        const T4<T>* __restrict__ U2, const T4<T>* __restrict__ PA,
                                                        const T4<T>* __restrict__ PB, T4<T>* __restrict__ O1,
                                                        T4<T>* __restrict__ O2, const T* clk) {
  constexpr int NFX = (WX + 1) * WY * WZ;
  constexpr int NFY = WX * (WY + 1) * WZ;
  constexpr int NFZ = WX * WY * (WZ + 1);
  constexpr int NF = NFX + NFY + NFZ;
  constexpr int NT = WX * WY * WZ;
  __shared__ T tf[8 * NF];
  const int x0 = blockIdx.x * WX, y0 = blockIdx.y * WY, z0 = blockIdx.z * WZ;
  const int li = threadIdx.x + WX * (threadIdx.y + WY * threadIdx.z);
  const int nfz = P.flat ? 0 : NFZ;
  for (int f = li; f < NFX + NFY + nfz; f += NT) {
    int axis, fx, fy, fz;
    if (f < NFX) {
      axis = 0; fx = f % (WX + 1); fy = (f / (WX + 1)) % WY; fz = f / ((WX + 1) * WY);
    } else if (f < NFX + NFY) {
      int g = f - NFX;
      axis = 1; fx = g % WX; fy = (g / WX) % (WY + 1); fz = g / (WX * (WY + 1));
    } else {
      int g = f - NFX - NFY;
      axis = 2; fx = g % WX; fy = (g / WX) % WY; fz = g / (WX * WY);
    }
    // the face between cell a and a + e (tile.wgsl:186-191)
    int ai = x0 + fx - (axis == 0), aj = y0 + fy - (axis == 1), ak = z0 + fz - (axis == 2);
    Flux<T> F = flux_between<T, LIM, RS, ISEN>(P, U1, U2, PA, PB, ai, aj, ak, axis);
    put_flux(tf, NF, f, F);
  }
  __syncthreads();
  int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  int i = x0 + tx, j = y0 + ty, k = z0 + tz;
  if (i >= P.n[0] || j >= P.n[1] || k >= P.n[2]) return;
  Flux<T> xl = get_flux(tf, NF, tx + (WX + 1) * (ty + WY * tz));
  Flux<T> xh = get_flux(tf, NF, tx + 1 + (WX + 1) * (ty + WY * tz));
  Flux<T> yl = get_flux(tf, NF, NFX + tx + WX * (ty + (WY + 1) * tz));
  Flux<T> yh = get_flux(tf, NF, NFX + tx + WX * (ty + 1 + (WY + 1) * tz));
  Flux<T> zl = zero_flux<T>(), zh = zero_flux<T>();
  if (!P.flat) {
    zl = get_flux(tf, NF, NFX + NFY + tx + WX * (ty + WY * tz));
    zh = get_flux(tf, NF, NFX + NFY + tx + WX * (ty + WY * (tz + 1)));
  }
  size_t o = cell_index(P, i, j, k);
  if constexpr (DEBUG) {
    // DEBUG_FLUX (hydro/update.wgsl:313-317)
    T4<T> a, b;
    a.x = xl.m; a.y = xh.m; a.z = yl.m; a.w = yh.m;
    b.x = zl.m; b.y = zh.m; b.z = T(0); b.w = T(0);
    O1[o] = a;
    O2[o] = b;
    return;
  }
  Flux<T> net = net_of(xl, xh, yl, yh, zl, zh);
  T dtdx = clk[0] / P.dx;
  T4<T> o1, o2;
  advance(P, U1[o], U2[o], net, PB[o].x, dtdx, o1, o2);
  O1[o] = o1;
  O2[o] = o2;
}

// ---------------- fused path ----------------
// LDS of the fused kernel, structure of arrays over the block's NT = BX x BY columns:
//   q[2][7]        primitives (rho, vx, vy, vz, p, s, c) in two alternating slots: plane c (the
//                  x/y neighbours and the column's own centre) and plane c - 1 (the column's z - 1
//                  neighbour), which each column then overwrites with its plane c + 1
//   sx[9], sy[9]   each column's upper-x and upper-y face side (SideData), for its neighbours
//   fx[8], fy[8]   the x and y fluxes of the face below each column; they reuse sx/sy's space
//                  once every face has read its sides (a barrier apart)
template <typename T, int NT>
struct FusedLayout {
  static constexpr int Q0 = 0, Q1 = 7 * NT, SX = 14 * NT, SY = 23 * NT, FX = 14 * NT, FY = 22 * NT;
  static constexpr int SIZE = 32 * NT;
};

template <typename T, int NT>
__device__ __forceinline__ Prim<T> lds_prim(const T* L, int t) {
  Prim<T> p;
  p.rho = L[0 * NT + t]; p.vx = L[1 * NT + t]; p.vy = L[2 * NT + t]; p.vz = L[3 * NT + t];
  p.p = L[4 * NT + t]; p.s = L[5 * NT + t];
  p.cs = T(0);
  return p;
}
template <typename T, int NT>
__device__ __forceinline__ void lds_put_prim(T* L, int t, const Prim<T>& q) {
  L[0 * NT + t] = q.rho; L[1 * NT + t] = q.vx; L[2 * NT + t] = q.vy; L[3 * NT + t] = q.vz;
  L[4 * NT + t] = q.p; L[5 * NT + t] = q.s; L[6 * NT + t] = q.cs;
}
template <typename T, int NT>
__device__ __forceinline__ Prim<T> lds_prim_c(const T* L, int t) {
  Prim<T> p = lds_prim<T, NT>(L, t);
  p.cs = L[6 * NT + t];
  return p;
}
template <typename T, int NT>
__device__ __forceinline__ void lds_put_side(T* L, int base, int t, const SideData<T>& d) {
  L[base + 0 * NT + t] = d.rho; L[base + 1 * NT + t] = d.vx; L[base + 2 * NT + t] = d.vy;
  L[base + 3 * NT + t] = d.vz; L[base + 4 * NT + t] = d.p; L[base + 5 * NT + t] = d.s;
  L[base + 6 * NT + t] = d.hrho; L[base + 7 * NT + t] = d.hp; L[base + 8 * NT + t] = d.shock;
}
template <typename T, int NT>
__device__ __forceinline__ SideData<T> lds_get_side(const T* L, int base, int t) {
  SideData<T> d;
  d.rho = L[base + 0 * NT + t]; d.vx = L[base + 1 * NT + t]; d.vy = L[base + 2 * NT + t];
  d.vz = L[base + 3 * NT + t]; d.p = L[base + 4 * NT + t]; d.s = L[base + 5 * NT + t];
  d.hrho = L[base + 6 * NT + t]; d.hp = L[base + 7 * NT + t]; d.shock = L[base + 8 * NT + t];
  return d;
}
template <typename T, int NT>
__device__ __forceinline__ void lds_put_flux(T* L, int base, int t, const Flux<T>& F) {
  L[base + 0 * NT + t] = F.m; L[base + 1 * NT + t] = F.mx; L[base + 2 * NT + t] = F.my; L[base + 3 * NT + t] = F.mz;
  L[base + 4 * NT + t] = F.e; L[base + 5 * NT + t] = F.ei; L[base + 6 * NT + t] = F.s; L[base + 7 * NT + t] = F.vn;
}
template <typename T, int NT>
__device__ __forceinline__ Flux<T> lds_get_flux(const T* L, int base, int t) {
  Flux<T> F;
  F.m = L[base + 0 * NT + t]; F.mx = L[base + 1 * NT + t]; F.my = L[base + 2 * NT + t]; F.mz = L[base + 3 * NT + t];
  F.e = L[base + 4 * NT + t]; F.ei = L[base + 5 * NT + t]; F.s = L[base + 6 * NT + t]; F.vn = L[base + 7 * NT + t];
  return F;
}


// With a non-periodic z face, a ghost plane (c < 0 or c >= nz) takes the half step of the cell it
// maps to (load_pred, hydro/update.wgsl:47-63), with that cell's own stencil: plane 0 or nz - 1,
// held in ZA/ZB (k_zplanes), with its normal velocity seen through the face. (Beyond a periodic
// face of such a grid too: the wrapped cell's stencil meets the other, non-periodic face.)
// Other planes keep h. The override is written field by field into h: a form that returns either
// h or the loaded state lets the compiler select between their addresses, which puts h in scratch
// memory on every plane.
template <typename T>
__device__ __forceinline__ void zghost_half(const Params<T>& P, const T4<T>* __restrict__ ZA,
                                            const T4<T>* __restrict__ ZB, int c, size_t col, Half<T>& h) {
  const int nz = P.n[2];
  if (c >= 0 && c < nz) return;
  const AxisMap mz = map_axis(c, nz, P.bc_lo[2], P.bc_hi[2]);
  const size_t o = (mz.i == 0 ? 0 : (size_t)P.n[0] * P.n[1]) + col;
  const T4<T> a = ZA[o], b = ZB[o];
  h.rho = a.x; h.vx = a.y; h.vy = a.z; h.vz = bc_vel(a.w, mz.flip, mz.side);
  h.p = b.x; h.s = b.y; h.shock = b.z; h.fo = b.w;
}

// The fused step. Block BX x BY threads; thread (tx, ty) owns column (x0 + tx - 2, y0 + ty - 2).
// Grid (ceil(nx / TX), ceil(ny / TY), ceil(nz / zc)). Planes [z0, z0 + zc) are updated.
// Boundary conditions along x and y: periodic, outflow or reflecting (an axis periodic on one
// face only goes to the plain path); along z: any, a non-periodic face through the boundary
// planes' half steps (ZA, ZB: zghost_half).
//
// Launch bound: in FP32 at least four waves per SIMD (two 512-thread blocks or one of 1024 per
// CU), which caps the kernel at 128 VGPRs; the FP64 twin, whose values take two registers each,
// is left uncapped.
//
// Per plane c of the march (4 barriers):
//   1. plane c's primitives (converted last iteration) into LDS                      | barrier
//   2. half step of plane c from the LDS x/y neighbours and the register z neighbours;
//      its upper x and y sides into LDS                                              | barrier
//   3. the x and y faces below this column and the z face below plane c             | barrier
//   4. those x/y fluxes into LDS (over the sides)                                    | barrier
//   5. plane c's x/y net inflow; the update of plane c - 1 (its state re-read, from L2) and
//      the next step's signal rate
template <typename T, int LIM, int RS, bool ISEN, int BX, int BY, bool DEBUG, bool MAPBC>
__global__ __launch_bounds__(BX* BY, (sizeof(T) == 4 ? 4 : 1)) void k_fused(
    Params<T> P, const T4<T>* __restrict__ U1, const T4<T>* __restrict__ U2, T4<T>* __restrict__ O1,
    T4<T>* __restrict__ O2, StepClock<T> C, const T4<T>* __restrict__ ZA, const T4<T>* __restrict__ ZB, int zc) {
  constexpr int NT = BX * BY;
  constexpr int TX = BX - 4, TY = BY - 4;
  using LY = FusedLayout<T, NT>;
  __shared__ T L[LY::SIZE];
  __shared__ T red[NT / 64 + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int t = tx + BX * ty;
  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  const int x = blockIdx.x * TX + tx - 2;
  const int y = blockIdx.y * TY + ty - 2;
  const int z0 = blockIdx.z * zc;
  const int z1 = min(z0 + zc, nz);
  const T dt = step_dt(P, C, 0);
  const T dtdx = dt / P.dx;
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && t == 0) step_clock_advance(P, C, 0);

  // this column's source through the x/y boundaries (clamped in range for the ragged rim,
  // whose results are never used)
  AxisMap mx = map_axis(x, nx, P.bc_lo[0], P.bc_hi[0]);
  AxisMap my = map_axis(y, ny, P.bc_lo[1], P.bc_hi[1]);
  const size_t col = (size_t)my.i * nx + mx.i;
  const size_t plane = (size_t)nx * ny;

  const bool in_r1 = tx >= 1 && tx <= BX - 2 && ty >= 1 && ty <= BY - 2;
  const bool real_x = x >= 0 && x < nx, real_y = y >= 0 && y < ny;
  const bool tile = tx >= 2 && tx <= BX - 3 && ty >= 2 && ty <= BY - 3 && real_x && real_y;
  const bool xface = tx >= 2 && tx <= BX - 2 && ty >= 2 && ty <= BY - 3 && x <= nx && real_y;
  const bool yface = ty >= 2 && ty <= BY - 2 && tx >= 2 && tx <= BX - 3 && y <= ny && real_x;
  // A ghost column beyond a non-periodic face takes the half step of the cell it maps to
  // (load_pred, hydro/update.wgsl:47-63): that cell's stencil sits one column inward.
  int sx = tx, sy = ty;
  if (!real_x && ((x < 0 && P.bc_lo[0] != BC_PERIODIC) || (x >= nx && P.bc_hi[0] != BC_PERIODIC))) sx = tx + (mx.i - x);
  if (!real_y && ((y < 0 && P.bc_lo[1] != BC_PERIODIC) || (y >= ny && P.bc_hi[1] != BC_PERIODIC))) sy = ty + (my.i - y);
  // (MAPBC: the grid has a non-periodic face; without one no column or plane is mapped and the
  // kernel is compiled without that path)
  const bool mapped = MAPBC && ((sx != tx) || (sy != ty));
  const int ts = min(max(sx, 1), BX - 2) + BX * min(max(sy, 1), BY - 2);

  // z: periodic, or (ZA given) a non-periodic face that ghost planes map through
  const bool zbc = MAPBC && ZA != nullptr;
  auto plane_of = [&](int z) {
    const int zi = zbc ? map_axis(z, nz, P.bc_lo[2], P.bc_hi[2]).i : ((z % nz) + nz) % nz;
    return (size_t)zi * plane + col;
  };
  auto prim_of = [&](T4<T> u1, T4<T> u2, int z) {
    Prim<T> q = prim_from_cons(P, u1, u2);
    q.vx = bc_vel(q.vx, mx.flip, mx.side);
    q.vy = bc_vel(q.vy, my.flip, my.side);
    if (zbc && (z < 0 || z >= nz)) {
      const AxisMap mz = map_axis(z, nz, P.bc_lo[2], P.bc_hi[2]);
      q.vz = bc_vel(q.vz, mz.flip, mz.side);
    }
    return q;
  };

  // prologue: planes z0 - 2 and z0 - 1 into the slots that iteration c = z0 - 1 reads as its
  // c - 1 and c; plane z0 in registers
  T4<T> b1, b2;
  size_t o0 = plane_of(z0 - 2);
  lds_put_prim<T, NT>(L + LY::Q0, t, prim_of(U1[o0], U2[o0], z0 - 2));
  o0 = plane_of(z0 - 1);
  lds_put_prim<T, NT>(L + LY::Q1, t, prim_of(U1[o0], U2[o0], z0 - 1));
  o0 = plane_of(z0);
  b1 = U1[o0]; b2 = U2[o0];

  SideData<T> zs_prev;                 // the upper z side of plane c - 1
  Flux<T> fz_lo = zero_flux<T>();      // the z face below plane c - 1
  Flux<T> axy_prev = zero_flux<T>();   // plane c - 1's (x_lo - x_hi) + (y_lo - y_hi)
  T rate_max = T(0);

  for (int c = z0 - 1; c <= z1; c++) {
    Prim<T> pp = prim_of(b1, b2, c + 1);
    if (c + 2 <= z1 + 1) {             // prefetch plane c + 2
      size_t o = plane_of(c + 2);
      b1 = U1[o];
      b2 = U2[o];
    }
    // 1. plane c is in slot Lq (written last iteration), plane c - 1 in Lm; the barrier also
    //    keeps this iteration's side writes behind the last iteration's flux reads (they share
    //    LDS)
    const bool odd = ((c - z0 + 1) & 1) != 0;
    const T* Lq = L + (odd ? LY::Q0 : LY::Q1);
    T* Lm = L + (odd ? LY::Q1 : LY::Q0);
    __syncthreads();

    // 2. half step and sides of plane c (hydro/predictor.wgsl:46-209)
    SideData<T> sxl, syl, szl, szu;
    if (in_r1) {
      Prim<T> pc = lds_prim_c<T, NT>(Lq, t);
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
        // the mapped cell's own stencil: its centre and z neighbours are this column's
        // primitives seen unflipped; its x/y neighbours are in LDS
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
      if (zbc && tile) zghost_half(P, ZA, ZB, c, (size_t)y * nx + x, hc);
      sxl = side_of(hc, ixd, T(-1));
      syl = side_of(hc, iyd, T(-1));
      szl = side_of(hc, izd, T(-1));
      szu = side_of(hc, izu, T(1));
      // what the neighbours above need
      lds_put_side<T, NT>(L, LY::SX, t, side_of(hc, ixu, T(1)));
      lds_put_side<T, NT>(L, LY::SY, t, side_of(hc, iyu, T(1)));
    }
    // plane c + 1 into the slot plane c - 1 leaves (only this column reads its own c - 1 entry)
    lds_put_prim<T, NT>(Lm, t, pp);
    __syncthreads();

    // 3. the x and y faces below this column, and the z face below plane c
    const bool live = c >= z0 && c < z1;
    Flux<T> Fx = zero_flux<T>(), Fy = zero_flux<T>(), fz = zero_flux<T>();
    if (live && xface) Fx = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SX, t - 1), sxl, 0);
    if (live && yface) Fy = face_from_sides<T, RS>(P, lds_get_side<T, NT>(L, LY::SY, t - BX), syl, 1);
    if (tile && c >= z0) fz = face_from_sides<T, RS>(P, zs_prev, szl, 2);
    __syncthreads();
    // 4. the x/y fluxes into LDS
    if (live && xface) lds_put_flux<T, NT>(L, LY::FX, t, Fx);
    if (live && yface) lds_put_flux<T, NT>(L, LY::FY, t, Fy);
    __syncthreads();

    // 5. this plane's x/y net inflow, and the update of plane c - 1
    Flux<T> axy = zero_flux<T>();
    if (live && tile) {
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
    if (tile && c - 1 >= z0 && c - 1 < z1) {
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
        // the next step's signal rate (hydro/rates.wgsl:60-62), from the state just made
        Prim<T> w = prim_from_cons(P, o1, o2);
        rate_max = tmax(rate_max, rate_of(P, w));
      }
      O1[o] = o1;
      O2[o] = o2;
    }
    // rotate
    zs_prev = szu;
    fz_lo = fz;
    axy_prev = axy;
  }
  if constexpr (!DEBUG) {
    // block maximum of the next step's rate (order-independent: max is exact)
    T v = wave_max(rate_max);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
      T m = T(0);
      for (int w = 0; w < (NT + 63) / 64; w++) m = tmax(m, red[w]);
      max_slot(C.next(0), (unsigned long long)order_bits(m));
    }
  }
}

}  // namespace shx
#include "uniform_roles.hpp"
namespace shx {

// ---------------- launchers ----------------
template <typename T, int LIM, int RS, bool ISEN>
struct UniformImpl {
  static void rates(const Launch& L, const void* u1, const void* u2, unsigned long long* slot) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    int blocks = (int)std::min<size_t>((L.ncell + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rates<T>, dim3(blocks), dim3(256), 0, L.stream, P, (const T4<T>*)u1, (const T4<T>*)u2,
                       (int)L.ncell, slot);
  }
  static void clock(const Launch& L, void* clk, unsigned long long* slot) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    hipLaunchKernelGGL(k_clock<T>, dim3(1), dim3(1), 0, L.stream, P, (T*)clk, slot);
  }
  static void predict(const Launch& L, const void* u1, const void* u2, void* pa, void* pb, const void* clk) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    dim3 b = P.flat ? dim3(16, 16, 1) : dim3(8, 8, 4);
    dim3 g((P.n[0] + b.x - 1) / b.x, (P.n[1] + b.y - 1) / b.y, (P.n[2] + b.z - 1) / b.z);
    hipLaunchKernelGGL((k_predict<T, LIM, ISEN>), g, b, 0, L.stream, P, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)pa,
                       (T4<T>*)pb, (const T*)clk);
  }
  static void update(const Launch& L, const void* u1, const void* u2, const void* pa, const void* pb, void* o1,
                     void* o2, const void* clk, bool debug) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    if (P.flat) {
      dim3 b(16, 16, 1), g((P.n[0] + 15) / 16, (P.n[1] + 15) / 16, P.n[2]);
      if (debug)
        hipLaunchKernelGGL((k_update<T, LIM, RS, ISEN, 16, 16, 1, true>), g, b, 0, L.stream, P, (const T4<T>*)u1,
                           (const T4<T>*)u2, (const T4<T>*)pa, (const T4<T>*)pb, (T4<T>*)o1, (T4<T>*)o2, (const T*)clk);
      else
        hipLaunchKernelGGL((k_update<T, LIM, RS, ISEN, 16, 16, 1, false>), g, b, 0, L.stream, P, (const T4<T>*)u1,
                           (const T4<T>*)u2, (const T4<T>*)pa, (const T4<T>*)pb, (T4<T>*)o1, (T4<T>*)o2, (const T*)clk);
    } else {
      dim3 b(8, 4, 4), g((P.n[0] + 7) / 8, (P.n[1] + 3) / 4, (P.n[2] + 3) / 4);
      if (debug)
        hipLaunchKernelGGL((k_update<T, LIM, RS, ISEN, 8, 4, 4, true>), g, b, 0, L.stream, P, (const T4<T>*)u1,
                           (const T4<T>*)u2, (const T4<T>*)pa, (const T4<T>*)pb, (T4<T>*)o1, (T4<T>*)o2, (const T*)clk);
      else
        hipLaunchKernelGGL((k_update<T, LIM, RS, ISEN, 8, 4, 4, false>), g, b, 0, L.stream, P, (const T4<T>*)u1,
                           (const T4<T>*)u2, (const T4<T>*)pa, (const T4<T>*)pb, (T4<T>*)o1, (T4<T>*)o2, (const T*)clk);
    }
  }
  template <int BX, int BY>
  static void fused_tile(const Launch& L, const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck,
                         const void* za, const void* zb, int zc, bool debug) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    constexpr int TX = BX - 4, TY = BY - 4;
    dim3 g((P.n[0] + TX - 1) / TX, (P.n[1] + TY - 1) / TY, (P.n[2] + zc - 1) / zc);
    bool mapbc = za != nullptr;
    for (int a = 0; a < 2; a++) mapbc |= P.bc_lo[a] != BC_PERIODIC || P.bc_hi[a] != BC_PERIODIC;
#define SHX_FUSED(D, M)                                                                                          \
  hipLaunchKernelGGL((k_fused<T, LIM, RS, ISEN, BX, BY, D, M>), g, dim3(BX, BY), 0, L.stream, P, (const T4<T>*)u1, \
                     (const T4<T>*)u2, (T4<T>*)o1, (T4<T>*)o2, step_clock<T>(ck), (const T4<T>*)za, (const T4<T>*)zb, zc)
    if (debug) { if (mapbc) SHX_FUSED(true, true); else SHX_FUSED(true, false); }
    else { if (mapbc) SHX_FUSED(false, true); else SHX_FUSED(false, false); }
#undef SHX_FUSED
  }
  template <int TX, int TY>
  static void fused_roles(const Launch& L, const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck,
                          const void* za, const void* zb, int zc, bool debug) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    dim3 g((P.n[0] + TX - 1) / TX, (P.n[1] + TY - 1) / TY, (P.n[2] + zc - 1) / zc);
    bool mapbc = za != nullptr;
    for (int a = 0; a < 2; a++) mapbc |= P.bc_lo[a] != BC_PERIODIC || P.bc_hi[a] != BC_PERIODIC;
#define SHX_FUSED_R(D, M)                                                                                            \
  hipLaunchKernelGGL((k_fused_r<T, LIM, RS, ISEN, TX, TY, D, M>), g, dim3(RoleTile<TX, TY>::NTH), 0, L.stream, P,     \
                     (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)o1, (T4<T>*)o2, step_clock<T>(ck), (const T4<T>*)za, \
                     (const T4<T>*)zb, zc)
    if (debug) { if (mapbc) SHX_FUSED_R(true, true); else SHX_FUSED_R(true, false); }
    else { if (mapbc) SHX_FUSED_R(false, true); else SHX_FUSED_R(false, false); }
#undef SHX_FUSED_R
  }
  // za, zb: the half steps of planes 0 and nz - 1 for a non-periodic z face (null: z periodic)
  static void fused(const Launch& L, int tile, const void* u1, const void* u2, void* o1, void* o2, const ClockRefs& ck,
                    const void* za, const void* zb, int zc, bool debug) {
    if constexpr (sizeof(T) == 8) {
      fused_tile<20, 20>(L, u1, u2, o1, o2, ck, za, zb, zc, debug);
    } else {
      switch (tile) {
        case FUSED_R32x16: fused_roles<32, 16>(L, u1, u2, o1, o2, ck, za, zb, zc, debug); break;
        case FUSED_R32x8: fused_roles<32, 8>(L, u1, u2, o1, o2, ck, za, zb, zc, debug); break;
        case FUSED_32x32: fused_tile<32, 32>(L, u1, u2, o1, o2, ck, za, zb, zc, debug); break;
        case FUSED_32x16: fused_tile<32, 16>(L, u1, u2, o1, o2, ck, za, zb, zc, debug); break;
        // 16 x 16 cells: threads by role in FP32
        default: fused_roles<16, 16>(L, u1, u2, o1, o2, ck, za, zb, zc, debug); break;
      }
    }
  }
  static void zplanes(const Launch& L, const void* u1, const void* u2, void* za, void* zb, const ClockRefs& ck) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    dim3 b(16, 16, 1), g((P.n[0] + 15) / 16, (P.n[1] + 15) / 16, 2);
    hipLaunchKernelGGL((k_zplanes<T, LIM, ISEN>), g, b, 0, L.stream, P, (const T4<T>*)u1, (const T4<T>*)u2, (T4<T>*)za,
                       (T4<T>*)zb, step_clock<T>(ck));
  }
  static void diag(const Launch& L, const DiagGeo& G, const void* u1, const void* u2, double* part, int nblocks,
                   double* out) {
    const Params<T>& P = *static_cast<const Params<T>*>(L.params);
    hipLaunchKernelGGL(k_diag<T>, dim3(nblocks), dim3(256), 0, L.stream, P, G, (const T4<T>*)u1, (const T4<T>*)u2,
                       (int)L.ncell, part);
    hipLaunchKernelGGL(k_diag_final<T>, dim3(1), dim3(32 * DIAG_CH), 0, L.stream, (const double*)part, nblocks, out);
  }
  static UniformOps ops() {
    UniformOps o;
    o.rates = rates;
    o.clock = clock;
    o.predict = predict;
    o.update = update;
    o.fused = fused;
    o.zplanes = zplanes;
    o.diag = diag;
    return o;
  }
};

}  // namespace shx
