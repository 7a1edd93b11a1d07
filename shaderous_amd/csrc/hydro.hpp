// Device core of the MUSCL-Hancock + HLLC hydro update: primitives, boundary mapping, limited
// reconstruction, the Hancock predictor, the face Riemann solve and the conservative update.
// Shared by every kernel (uniform plain, uniform fused, blocks, ensemble) so that all of them
// compute each quantity with the same instructions. Templated on the arithmetic type T (float
// for the production path, double for the all-FP64 precision twin), the limiter, the Riemann
// solver and the density-along-the-adiabat option.
//
// The arithmetic follows the reference's WGSL with every optional module off; each function
// cites the file and lines it follows (paths under the reference's src/).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shx {

enum { LIM_KOREN = 0, LIM_MC = 1, LIM_VANLEER = 2, LIM_MINMOD = 3 };
enum { RS_HLLC = 0, RS_HLL = 1 };
enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2, BC_OUTFLOW_ONLY = 3 };

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
template <typename T> using T4 = typename Vec4<T>::type;

// The Params values the hydro path reads (hydro/common.wgsl:4-79), in the run's precision.
template <typename T>
struct Params {
  int n[3];
  int bc_lo[3];
  int bc_hi[3];
  int flat;            // one layer along z (2D): z neighbours are the cell itself
  T ndim;              // P.geom.y
  T gamma, gm1, igm1;  // P.eos.x, P.eos.x - 1, 1 / (P.eos.x - 1)
  T dx, cfl, rho_floor, p_floor, eta_dual, shock_thresh, t_end, dt_max;
  T fluff[3];
};

// min/max/abs as single instructions (v_min_f32, v_max_f32, the |x| source modifier); a NaN
// operand yields the other one.
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float tabs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double tabs(double a) { return __builtin_fabs(a); }
template <typename T> __device__ __forceinline__ T tsign(T a) { return a > T(0) ? T(1) : (a < T(0) ? T(-1) : T(0)); }
template <typename T> __device__ __forceinline__ T tclamp(T x, T lo, T hi) { return tmin(tmax(x, lo), hi); }
// Single precision: the hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32, ~1 ulp;
// WGSL allows 2.5 ulp for division). Double precision: correctly rounded.
__device__ __forceinline__ float trcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double trcp(double b) { return 1.0 / b; }
__device__ __forceinline__ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double tdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
// Single-precision pow as WGSL defines it, exp2(b log2 a) (every base here is positive), by the
// hardware's log2 and exp2 (v_log_f32, v_exp_f32): the library powf's exact-case handling costs
// ~40 instructions a call, and the isentropic reconstruction makes six calls per axis.
__device__ __forceinline__ float tpow(float a, float b) { return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a)); }
__device__ __forceinline__ double tpow(double a, double b) { return pow(a, b); }
// WGSL mix(x, y, a) = x (1 - a) + y a
template <typename T> __device__ __forceinline__ T tmix(T x, T y, T a) { return x * (T(1) - a) + y * a; }
// sign(a) * sign(b) > 0, formed from the signs so that small products do not underflow
// (limiters/koren.wgsl:287-289)
template <typename T> __device__ __forceinline__ bool same_sign(T a, T b) {
  return (a > T(0) && b > T(0)) || (a < T(0) && b < T(0));
}
// minmod(x, y) = sign(x) min(|x|, |y|) when sign(x) sign(y) > 0, else 0: in FP32 the median of
// (0, x, y), one instruction (a zero result may carry the sign of x or y)
__device__ __forceinline__ float minmod(float x, float y) { return __builtin_amdgcn_fmed3f(0.0f, x, y); }
__device__ __forceinline__ double minmod(double x, double y) {
  return same_sign(x, y) ? (x > 0.0 ? tmin(x, y) : tmax(x, y)) : 0.0;
}

// Primitive record (hydro/q.wgsl:56): rho, v, p, scalar s, sound speed c.
template <typename T>
struct Prim {
  T rho, vx, vy, vz, p, s, cs;
};

// prim_from_cons (hydro/prim.wgsl:131-156), ideal gas (eos/ideal.wgsl:235-267), pressure floor
// without a Jeans floor (hydro/pfloor.wgsl:169-171).
template <typename T>
__device__ __forceinline__ Prim<T> prim_from_cons(const Params<T>& P, T4<T> u1, T4<T> u2) {
  Prim<T> q;
  T rho = tmax(u1.x, P.rho_floor);
  T p = P.gm1 * u2.y;
  if (!(p >= P.p_floor)) p = P.p_floor;
  T ir = trcp(rho);
  q.rho = rho;
  q.vx = u1.y * ir;
  q.vy = u1.z * ir;
  q.vz = u1.w * ir;
  q.p = p;
  q.s = u2.z * ir;
  q.cs = tsqrt(P.gamma * p * ir);
  return q;
}

// ---- boundary mapping: hydro/common.wgsl:433-539, hydro/boundary.wgsl:47-84 ----
// One axis: the in-range index standing for i, the velocity sign along that axis and the
// outflow-only side (+-1, 0 for none).
struct AxisMap { int i; int flip; int side; };
__device__ __forceinline__ AxisMap map_axis(int i, int n, int lo, int hi) {
  AxisMap m{i, 1, 0};
  if (i >= 0 && i < n) return m;
  bool up = i >= n;
  int k = up ? hi : lo;
  if (k == BC_PERIODIC) {
    int v = up ? i - n : i + n;
    m.i = v < 0 ? 0 : (v > n - 1 ? n - 1 : v);
  } else if (k == BC_REFLECT) {
    m.i = up ? 2 * n - 1 - i : -1 - i;
    m.flip = -1;
  } else {
    m.i = up ? n - 1 : 0;
    if (k == BC_OUTFLOW_ONLY) m.side = up ? 1 : -1;
  }
  // (a reflecting axis of one or two cells, or a ragged tile's rim, can reach further than n
  // beyond the face: keep every index in range)
  m.i = m.i < 0 ? 0 : (m.i > n - 1 ? n - 1 : m.i);
  return m;
}
// bc_velocity (common.wgsl:536-539) on one component
template <typename T>
__device__ __forceinline__ T bc_vel(T v, int flip, int side) {
  T w = flip < 0 ? -v : v;
  if (side != 0) {
    T sd = T(side);
    w = tmax(w * sd, T(0)) * sd;
  }
  return w;
}

// ---- limiters: hydro/limiters/*.wgsl, per channel ----
// Koren's smoothness measure of a channel against its natural scale s = 0.01 scale: the
// difference energy over s^2 (koren.wgsl:274-292), or, where s^2 over- or underflows, the
// differences scaled by 1/|s| first. What depends on the scale alone is formed once per scale
// (a cell's six channels have four: rho, c_s three times, p, none) and shared by its edges.
template <typename T>
struct EdgeScale { T s2inv; T k; bool as_is; };

template <typename T>
__device__ __forceinline__ EdgeScale<T> edge_scale(T scale) {
  const T SMOOTH = T(0.01);
  T s = SMOOTH * scale;
  T s2 = s * s;
  EdgeScale<T> e;
  e.as_is = (s2 >= T(1e-30) && s2 < T(1e30)) || (s == T(0));
  e.s2inv = tmax(s2, T(1e-30));
  if constexpr (sizeof(T) == 4) e.s2inv = trcp(e.s2inv);   // tdiv(q, ss) = q rcp(ss) in FP32
  e.k = trcp(tmax(tabs(s), T(1e-37)));
  return e;
}

// Twice the edge increment from the differences (far, near) and the channel's scale.
// Koren (koren.wgsl:274-292): its early return gives exactly the blend's value per channel
// (the blend weight is 1 for a smooth channel), so each channel is formed on its own.
template <typename T, int LIM>
__device__ __forceinline__ T edge1(T far, T near, const EdgeScale<T>& sc) {
  if constexpr (LIM == LIM_KOREN) {
    T fr = (far + T(2) * near) * T(1.0 / 3.0);
    T q = far * far + near * near;
    T eta;
    if (sc.as_is) {
      // FP32: q times the reciprocal (tdiv's form); FP64: the quotient (s2inv holds ss)
      if constexpr (sizeof(T) == 4) eta = q * sc.s2inv;
      else eta = tdiv(q, sc.s2inv);
    } else {
      T fk = far * sc.k, nk = near * sc.k;
      eta = fk * fk + nk * nk;
    }
    // sign(near) min(2|far|, 2|near|, |fr|) where far and near share a sign (fr then does too)
    T lim = minmod(T(2) * far, minmod(T(2) * near, fr));
    T w = tclamp(T(1.5) - eta, T(0), T(1));
    return (w == T(1)) ? fr : tmix(lim, fr, w);
  } else if constexpr (LIM == LIM_MC) {
    // sign(far) min(2|far|, 2|near|, |far + near| / 2) where far and near share a sign
    return minmod(T(2) * far, minmod(T(2) * near, T(0.5) * (far + near)));
  } else if constexpr (LIM == LIM_MINMOD) {
    return minmod(far, near);
  } else {
    return same_sign(far, near) ? T(2) * far * tdiv(near, far + near) : T(0);
  }
}
// The same from the scale itself (callers with one channel per scale)
template <typename T, int LIM>
__device__ __forceinline__ T edge1(T far, T near, T scale) {
  EdgeScale<T> sc{};
  if constexpr (LIM == LIM_KOREN) sc = edge_scale(scale);
  return edge1<T, LIM>(far, near, sc);
}
template <int LIM> constexpr bool lim_symmetric() { return LIM != LIM_KOREN; }

// Twice the increment of each reconstructed channel (rho, vx, vy, vz, p, s) toward a face.
template <typename T>
struct Inc { T rho, vx, vy, vz, p, s; };

// isen_inc / isen_blend (hydro/q.wgsl:183-213) with no reference state and no potential, in two
// parts: what depends only on the three cells (the densities along the adiabat through the
// centre, the smoothness, adiabat and shear switches), shared by a cell's up and down edges, and
// the edge's own blend.
template <typename T>
struct IsenCommon { bool ok; T gi, tl, tr, room; };

template <typename T>
__device__ __forceinline__ IsenCommon<T> isen_common(const Params<T>& P, const Prim<T>& l, const Prim<T>& c,
                                                     const Prim<T>& r) {
  const T ISEN_SMOOTH = T(0.2), ISEN_ADIABAT = T(0.01), ISEN_SHEAR = T(0.1), ISEN_ROOM = T(0.01);
  IsenCommon<T> k;
  k.ok = l.rho > T(0) && l.p > T(0) && c.rho > T(0) && c.p > T(0) && r.rho > T(0) && r.p > T(0);
  k.gi = trcp(P.gamma);
  k.tl = l.rho * tpow(tdiv(c.p, l.p), k.gi) - c.rho;
  k.tr = r.rho * tpow(tdiv(c.p, r.p), k.gi) - c.rho;
  T ks = tclamp(T(2) - tdiv(tmax(tabs(l.p - c.p), tabs(r.p - c.p)), ISEN_SMOOTH * c.p), T(0), T(1));
  T ka = tclamp(T(2) - tdiv(tmax(tabs(k.tl), tabs(k.tr)), ISEN_ADIABAT * c.rho), T(0), T(1));
  T dxr = r.vx - c.vx, dyr = r.vy - c.vy, dzr = r.vz - c.vz;
  T dxl = c.vx - l.vx, dyl = c.vy - l.vy, dzl = c.vz - l.vz;
  T dvr = tsqrt(dxr * dxr + dyr * dyr + dzr * dzr);
  T dvl = tsqrt(dxl * dxl + dyl * dyl + dzl * dzl);
  T kv = tclamp(T(2) - tdiv(tmax(dvr, dvl), ISEN_SHEAR * tmax(c.cs, T(1e-30))), T(0), T(1));
  k.room = ISEN_ROOM * c.rho * ks * ka * kv;
  return k;
}

template <typename T, int LIM>
__device__ __forceinline__ T isen_blend(const IsenCommon<T>& k, const EdgeScale<T>& srho, const Prim<T>& l,
                                        const Prim<T>& c, const Prim<T>& r, T x, T bx, T sgn) {
  const T ISEN_CAP = T(0.25);
  if (!k.ok) return x + tclamp(T(0) - x, T(0), T(0));
  T far = sgn > T(0) ? -k.tl : k.tr;
  T near = sgn > T(0) ? k.tr : -k.tl;
  T d = edge1<T, LIM>(far, near, srho);
  T pf = c.p + T(0.5) * sgn * bx;
  T wn = sgn > T(0) ? r.rho : l.rho;
  T rf = tclamp((c.rho + T(0.5) * sgn * d) * tpow(tdiv(tmax(pf, T(1e-30)), c.p), k.gi), tmin(c.rho, wn),
                tmax(c.rho, wn));
  T bound = tmax(ISEN_CAP * tmin(tabs(c.rho - l.rho), tabs(r.rho - c.rho)), k.room);
  T i0 = T(2) * sgn * (rf - c.rho);
  return x + tclamp(i0 - x, -bound, bound);
}

// edge_up (UP) / edge_dn (hydro/q.wgsl:214-239), no potential.
template <typename T, int LIM, bool ISEN, bool UP>
__device__ __forceinline__ Inc<T> edge_k(const Params<T>& P, const Prim<T>& l, const Prim<T>& c, const Prim<T>& r,
                                         const IsenCommon<T>& k) {
  Inc<T> o;
  constexpr bool fwd = UP || lim_symmetric<LIM>();
  auto e = [&](T lo, T mid, T hi, const EdgeScale<T>& sc) {
    T below = mid - lo, above = hi - mid;
    return fwd ? edge1<T, LIM>(below, above, sc) : edge1<T, LIM>(above, below, sc);
  };
  EdgeScale<T> srho{}, sv{}, sp{}, s0{};
  if constexpr (LIM == LIM_KOREN) {
    srho = edge_scale(c.rho); sv = edge_scale(c.cs); sp = edge_scale(c.p); s0 = edge_scale(T(0));
  }
  o.rho = e(l.rho, c.rho, r.rho, srho);
  o.vx = e(l.vx, c.vx, r.vx, sv);
  o.vy = e(l.vy, c.vy, r.vy, sv);
  o.vz = e(l.vz, c.vz, r.vz, sv);
  o.p = e(l.p, c.p, r.p, sp);
  o.s = e(l.s, c.s, r.s, s0);
  if constexpr (ISEN) o.rho = isen_blend<T, LIM>(k, srho, l, c, r, o.rho, o.p, UP ? T(1) : T(-1));
  return o;
}
template <typename T, int LIM, bool ISEN, bool UP>
__device__ __forceinline__ Inc<T> edge(const Params<T>& P, const Prim<T>& l, const Prim<T>& c, const Prim<T>& r) {
  IsenCommon<T> k{};
  if constexpr (ISEN) k = isen_common(P, l, c, r);
  return edge_k<T, LIM, ISEN, UP>(P, l, c, r, k);
}

// Half-step state (hydro/predictor.wgsl:46-209): rho, v, p, s at t + dt/2, the shock flag and
// the first-order flag.
template <typename T>
struct Half { T rho, vx, vy, vz, p, s, shock, fo; };

// One axis's contribution to the predictor: the face-to-face slope and the positivity terms.
template <typename T>
struct AxisSlope { Inc<T> s; T big; T fall; };

template <typename T, int LIM, bool ISEN>
__device__ __forceinline__ AxisSlope<T> axis_slope(const Params<T>& P, const Prim<T>& m, const Prim<T>& q,
                                                   const Prim<T>& p, Inc<T>* up_out, Inc<T>* dn_out) {
  AxisSlope<T> r;
  // a cell's up and down edges share the isentropic terms of its three cells
  IsenCommon<T> k{};
  if constexpr (ISEN) k = isen_common(P, m, q, p);
  Inc<T> u = edge_k<T, LIM, ISEN, true>(P, m, q, p, k);
  if constexpr (lim_symmetric<LIM>()) {
    r.s = u;
    r.big = tabs(u.p);
    r.fall = T(-0.5) * tabs(u.rho);
    if (up_out) *up_out = u;
    if (dn_out) *dn_out = u;
  } else {
    Inc<T> d = edge_k<T, LIM, ISEN, false>(P, m, q, p, k);
    r.s.rho = T(0.5) * (u.rho + d.rho);
    r.s.vx = T(0.5) * (u.vx + d.vx);
    r.s.vy = T(0.5) * (u.vy + d.vy);
    r.s.vz = T(0.5) * (u.vz + d.vz);
    r.s.p = T(0.5) * (u.p + d.p);
    r.s.s = T(0.5) * (u.s + d.s);
    r.big = tmax(tabs(u.p), tabs(d.p));
    r.fall = tmin(T(0.5) * u.rho, T(-0.5) * d.rho);
    if (up_out) *up_out = u;
    if (dn_out) *dn_out = d;
  }
  return r;
}

// The Hancock half step from the cell, its six neighbours' pressures and normal velocities
// (for the shock detector) and the three axis slopes.
template <typename T>
__device__ __forceinline__ Half<T> hancock(const Params<T>& P, const Prim<T>& q, const AxisSlope<T> sl[3],
                                           T conv, T jump, T dt) {
  Half<T> o;
  T big = tmax(tmax(tmax(T(0), sl[0].big), sl[1].big), sl[2].big);
  T fall = tmin(tmin(tmin(T(0), sl[0].fall), sl[1].fall), sl[2].fall);
  T shock = (conv < T(0) && jump > P.shock_thresh) ? T(1) : T(0);
  T h = T(0.5) * dt / P.dx;
  T rho = q.rho;
  const Inc<T>& sx = sl[0].s;
  const Inc<T>& sy = sl[1].s;
  const Inc<T>& sz = sl[2].s;
  T dv = sx.vx + sy.vy + sz.vz;
  T da0 = q.vx * sx.rho + q.vy * sy.rho + q.vz * sz.rho;
  T da1 = q.vx * sx.vx + q.vy * sy.vx + q.vz * sz.vx;
  T da2 = q.vx * sx.vy + q.vy * sy.vy + q.vz * sz.vy;
  T da3 = q.vx * sx.vz + q.vy * sy.vz + q.vz * sz.vz;
  T db0 = q.vx * sx.p + q.vy * sy.p + q.vz * sz.p;
  T db1 = q.vx * sx.s + q.vy * sy.s + q.vz * sz.s;
  T ir = trcp(rho);
  o.rho = rho - h * (da0 + rho * dv);
  o.vx = q.vx - h * (da1 + sx.p * ir);
  o.vy = q.vy - h * (da2 + sy.p * ir);
  o.vz = q.vz - h * (da3 + sz.p * ir);
  o.p = q.p - h * (db0 + rho * q.cs * q.cs * dv);
  o.s = q.s - h * db1;
  o.shock = shock;
  T rmin = o.rho + fall;
  T pmin = o.p - T(0.5) * big;
  o.fo = T(0);
  if (!(rmin > T(0) && pmin > T(0))) {
    o.rho = q.rho; o.vx = q.vx; o.vy = q.vy; o.vz = q.vz; o.p = q.p; o.s = q.s;
    o.fo = T(1);
  }
  return o;
}

// Shock detector (predictor.wgsl:114-134): converging flow and the largest relative pressure
// jump across the cell.
template <typename T>
__device__ __forceinline__ T jump_of(T pm, T pp) {
  return tdiv(tabs(pp - pm), tmin(pp, pm));
}

// The half step of a cell of a flat grid (one layer along z) from its four neighbours: hancock
// without the z axis, whose neighbours are the cell itself (predictor.wgsl:46-209 with a zero
// z slope, no z convergence and no z jump). The terms left out are zeros, so the result is
// hancock's but for the sign of a zero. ixu .. iyd (may be null) receive the x and y edge
// increments.
template <typename T, int LIM, bool ISEN>
__device__ __forceinline__ Half<T> half_flat(const Params<T>& P, const Prim<T>& q, const Prim<T>& xm,
                                             const Prim<T>& xp, const Prim<T>& ym, const Prim<T>& yp, T dt,
                                             Inc<T>* ixu, Inc<T>* ixd, Inc<T>* iyu, Inc<T>* iyd) {
  const AxisSlope<T> ax = axis_slope<T, LIM, ISEN>(P, xm, q, xp, ixu, ixd);
  const AxisSlope<T> ay = axis_slope<T, LIM, ISEN>(P, ym, q, yp, iyu, iyd);
  const T conv = (xp.vx - xm.vx) + (yp.vy - ym.vy);
  const T jump = tmax(jump_of(xm.p, xp.p), jump_of(ym.p, yp.p));
  Half<T> o;
  T big = tmax(tmax(T(0), ax.big), ay.big);
  T fall = tmin(tmin(T(0), ax.fall), ay.fall);
  T shock = (conv < T(0) && jump > P.shock_thresh) ? T(1) : T(0);
  T h = T(0.5) * dt / P.dx;
  T rho = q.rho;
  const Inc<T>& sx = ax.s;
  const Inc<T>& sy = ay.s;
  T dv = sx.vx + sy.vy;
  T da0 = q.vx * sx.rho + q.vy * sy.rho;
  T da1 = q.vx * sx.vx + q.vy * sy.vx;
  T da2 = q.vx * sx.vy + q.vy * sy.vy;
  T da3 = q.vx * sx.vz + q.vy * sy.vz;
  T db0 = q.vx * sx.p + q.vy * sy.p;
  T db1 = q.vx * sx.s + q.vy * sy.s;
  T ir = trcp(rho);
  o.rho = rho - h * (da0 + rho * dv);
  o.vx = q.vx - h * (da1 + sx.p * ir);
  o.vy = q.vy - h * (da2 + sy.p * ir);
  o.vz = q.vz - h * da3;
  o.p = q.p - h * (db0 + rho * q.cs * q.cs * dv);
  o.s = q.s - h * db1;
  o.shock = shock;
  T rmin = o.rho + fall;
  T pmin = o.p - T(0.5) * big;
  o.fo = T(0);
  if (!(rmin > T(0) && pmin > T(0))) {
    o.rho = q.rho; o.vx = q.vx; o.vy = q.vy; o.vz = q.vz; o.p = q.p; o.s = q.s;
    o.fo = T(1);
  }
  return o;
}

// ---- face: hydro/flux.wgsl:60-125, riemann/hllc.wgsl, riemann/hll.wgsl ----
template <typename T>
struct Flux { T m, mx, my, mz, e, ei, s, vn; };

// One side of a face (to_side, flux.wgsl:60-85) in the face frame: n normal, t1/t2 the
// cyclically following components (rot, flux.wgsl:37-41).
template <typename T>
struct Side { T rho, vn, vt1, vt2, p, s, cs, eint; };

template <typename T>
__device__ __forceinline__ Side<T> make_side(const Params<T>& P, T rho, T vn, T vt1, T vt2, T p, T s, T lo_rho, T lo_p) {
  Side<T> o;
  o.rho = tmax(rho, tmax(lo_rho, P.rho_floor));
  o.vn = vn; o.vt1 = vt1; o.vt2 = vt2;
  o.s = s;
  o.p = tmax(p, tmax(lo_p, P.p_floor));
  o.cs = tsqrt(tdiv(P.gamma * o.p, o.rho));
  o.eint = o.p * P.igm1;
  return o;
}

// HLLC blended toward HLL by w (riemann/hllc.wgsl:4-85), or HLL (riemann/hll.wgsl:87-121).
// A side's total energy, momentum and normal flux (hllc.wgsl:7-28), each from the side's own
// primitives in one fixed expression
template <typename T>
struct SideFlux { T E, m0, m1, m2, f0, f1, f2, fe; };
template <typename T>
__device__ __forceinline__ SideFlux<T> side_flux(T rho, T vn, T vt1, T vt2, T p, T eint) {
  SideFlux<T> o;
  o.E = eint + T(0.5) * rho * (vn * vn + vt1 * vt1 + vt2 * vt2);
  o.m0 = rho * vn; o.m1 = rho * vt1; o.m2 = rho * vt2;
  o.f0 = o.m0 * vn + p; o.f1 = o.m1 * vn; o.f2 = o.m2 * vn;
  o.fe = (o.E + p) * vn;
  return o;
}

// HLLC blended toward HLL by w (riemann/hllc.wgsl:4-85), or HLL (riemann/hll.wgsl:87-121).
// Each branch forms only the side quantities it uses: the upwind side's on a supersonic face,
// the star side's on a subsonic one, both where the HLL state is needed.
template <typename T, int RS>
__device__ __forceinline__ Flux<T> riemann(const Side<T>& l, const Side<T>& r, T w_hll) {
  T sl = tmin(l.vn - l.cs, r.vn - r.cs);
  T sr = tmax(l.vn + l.cs, r.vn + r.cs);
  Flux<T> o;
  if (sl >= T(0) || sr <= T(0)) {
    // supersonic: the upwind side's own flux
    const bool up_l = sl >= T(0);
    const SideFlux<T> u = side_flux(up_l ? l.rho : r.rho, up_l ? l.vn : r.vn, up_l ? l.vt1 : r.vt1,
                                    up_l ? l.vt2 : r.vt2, up_l ? l.p : r.p, up_l ? l.eint : r.eint);
    o.m = u.m0; o.mx = u.f0; o.my = u.f1; o.mz = u.f2; o.e = u.fe; o.vn = up_l ? l.vn : r.vn;
  } else {
    T ql = l.rho * (sl - l.vn);
    T qr = r.rho * (sr - r.vn);
    const bool star = RS != RS_HLL && ql < T(0) && qr > T(0);
    // The HLL state between the waves: the flux itself where the star densities are not
    // positive (and for HLL), the blend's other end where the face is flagged; a subsonic
    // unflagged face (w = 0: the blend is the star flux exactly) skips it.
    T hm, hp0, hp1, hp2, he;
    if (!star || w_hll != T(0)) {
      const SideFlux<T> a = side_flux(l.rho, l.vn, l.vt1, l.vt2, l.p, l.eint);
      const SideFlux<T> b = side_flux(r.rho, r.vn, r.vt1, r.vt2, r.p, r.eint);
      T inv = trcp(sr - sl);
      T slsr = sl * sr;
      hm = (sr * a.m0 - sl * b.m0 + slsr * (r.rho - l.rho)) * inv;
      hp0 = (sr * a.f0 - sl * b.f0 + slsr * (b.m0 - a.m0)) * inv;
      hp1 = (sr * a.f1 - sl * b.f1 + slsr * (b.m1 - a.m1)) * inv;
      hp2 = (sr * a.f2 - sl * b.f2 + slsr * (b.m2 - a.m2)) * inv;
      he = (sr * a.fe - sl * b.fe + slsr * (b.E - a.E)) * inv;
    }
    if constexpr (RS == RS_HLL) {
      o.m = hm; o.mx = hp0; o.my = hp1; o.mz = hp2; o.e = he;
      o.vn = tdiv(r.p - l.p + l.vn * ql - r.vn * qr, ql - qr);
    } else {
      if (!star) {
        o.m = hm; o.mx = hp0; o.my = hp1; o.mz = hp2; o.e = he;
        o.vn = T(0.5) * (l.vn + r.vn);
      } else {
        T sm = tdiv(r.p - l.p + l.vn * ql - r.vn * qr, ql - qr);
        // the star state as differences from the outer state (hllc.wgsl:52-74)
        bool left = sm >= T(0);
        T S = left ? sl : sr;
        T rho = left ? l.rho : r.rho;
        T un = left ? l.vn : r.vn;
        T u1 = left ? l.vt1 : r.vt1;
        T u2 = left ? l.vt2 : r.vt2;
        T p = left ? l.p : r.p;
        T q = left ? ql : qr;
        const SideFlux<T> c = side_flux(rho, un, u1, u2, p, left ? l.eint : r.eint);
        T k = tdiv(sm - un, S - sm);
        T dr = rho * k;
        T f = rho + dr;
        T cm = c.m0 + S * dr;
        T c0 = c.f0 + S * (sm * dr + rho * (sm - un));
        T c1 = c.f1 + S * (dr * u1);
        T c2 = c.f2 + S * (dr * u2);
        T ce = c.fe + S * (c.E * k + f * (sm - un) * (sm + tdiv(p, q)));
        if (w_hll != T(0)) {
          o.m = tmix(cm, hm, w_hll);
          o.mx = tmix(c0, hp0, w_hll);
          o.my = tmix(c1, hp1, w_hll);
          o.mz = tmix(c2, hp2, w_hll);
          o.e = tmix(ce, he, w_hll);
        } else {
          o.m = cm; o.mx = c0; o.my = c1; o.mz = c2; o.e = ce;
        }
        o.vn = sm;
      }
    }
  }
  // internal energy and the scalar ride on the mass flux, upwinded
  bool up = o.m >= T(0);
  o.ei = o.m * tdiv(up ? l.eint : r.eint, up ? l.rho : r.rho);
  o.s = o.m * (up ? l.s : r.s);
  return o;
}

// One cell's side of a face before the face's own lower bound and floors: the half-step state
// plus half the edge increment toward the face (to_side, flux.wgsl:60-63; zero increment for a
// first-order cell), with the cell's half-step density and pressure (for the face's lower bound,
// flux.wgsl:109-117) and its shock flag (flux.wgsl:93).
template <typename T>
struct SideData { T rho, vx, vy, vz, p, s, hrho, hp, shock; };

// sgn = +1 for the cell's upper face (il its up increment), -1 for its lower face.
template <typename T>
__device__ __forceinline__ SideData<T> side_of(const Half<T>& h, const Inc<T>& inc, T sgn) {
  T hs = T(0.5) * sgn * (T(1) - h.fo);
  SideData<T> o;
  o.rho = h.rho + hs * inc.rho;
  o.vx = h.vx + hs * inc.vx;
  o.vy = h.vy + hs * inc.vy;
  o.vz = h.vz + hs * inc.vz;
  o.p = h.p + hs * inc.p;
  o.s = h.s + hs * inc.s;
  o.hrho = h.rho;
  o.hp = h.p;
  o.shock = h.shock;
  return o;
}

// face_flux (flux.wgsl:92-125) from the left cell's upper side and the right cell's lower side
// along `axis`. Momentum comes back in grid components.
template <typename T, int RS>
__device__ __forceinline__ Flux<T> face_from_sides(const Params<T>& P, const SideData<T>& L, const SideData<T>& R,
                                                   int axis) {
  T w = tmax(L.shock, R.shock);
  T lo_rho = T(0.5) * tmin(L.hrho, R.hrho);
  T lo_p = T(0.5) * tmin(L.hp, R.hp);
  // rot (flux.wgsl:37-41): axis 0 -> (x, y, z), 1 -> (y, z, x), 2 -> (z, x, y)
  T ln, l1, l2, rn, r1, r2;
  if (axis == 0) { ln = L.vx; l1 = L.vy; l2 = L.vz; rn = R.vx; r1 = R.vy; r2 = R.vz; }
  else if (axis == 1) { ln = L.vy; l1 = L.vz; l2 = L.vx; rn = R.vy; r1 = R.vz; r2 = R.vx; }
  else { ln = L.vz; l1 = L.vx; l2 = L.vy; rn = R.vz; r1 = R.vx; r2 = R.vy; }
  Side<T> sl = make_side(P, L.rho, ln, l1, l2, L.p, L.s, lo_rho, lo_p);
  Side<T> sr = make_side(P, R.rho, rn, r1, r2, R.p, R.s, lo_rho, lo_p);
  Flux<T> F = riemann<T, RS>(sl, sr, w);
  // unrot: the face frame's (n, t1, t2) back to grid components
  T n = F.mx, t1 = F.my, t2 = F.mz;
  if (axis == 0) { F.mx = n; F.my = t1; F.mz = t2; }
  else if (axis == 1) { F.mx = t2; F.my = n; F.mz = t1; }
  else { F.mx = t1; F.my = t2; F.mz = n; }
  return F;
}

// face_from_sides for an axis known only at run time: the velocities are rotated into the face
// frame here and the face solved as one along x, the same values in the same order as
// face_from_sides(axis). The components are picked from registers (the empty asm makes them
// opaque values): picking them from the side structs lets the compiler select by address, which
// puts both sides in scratch memory.
template <typename T, int RS>
__device__ __forceinline__ Flux<T> face_from_sides_rt(const Params<T>& P, SideData<T> L, SideData<T> R, int axis) {
  T lx = L.vx, ly = L.vy, lz = L.vz, rx = R.vx, ry = R.vy, rz = R.vz;
  asm("" : "+v"(lx), "+v"(ly), "+v"(lz), "+v"(rx), "+v"(ry), "+v"(rz));
  const bool a0 = axis == 0, a1 = axis == 1;
  L.vx = a0 ? lx : (a1 ? ly : lz); L.vy = a0 ? ly : (a1 ? lz : lx); L.vz = a0 ? lz : (a1 ? lx : ly);
  R.vx = a0 ? rx : (a1 ? ry : rz); R.vy = a0 ? ry : (a1 ? rz : rx); R.vz = a0 ? rz : (a1 ? rx : ry);
  Flux<T> F = face_from_sides<T, RS>(P, L, R, 0);
  T n = F.mx, t1 = F.my, t2 = F.mz;
  asm("" : "+v"(n), "+v"(t1), "+v"(t2));
  F.mx = a0 ? n : (a1 ? t2 : t1); F.my = a0 ? t1 : (a1 ? n : t2); F.mz = a0 ? t2 : (a1 ? t1 : n);
  return F;
}

// face_flux between the half-step states hl (left) and hr (right), given twice the increments
// from each cell toward the face (il: the left cell's up increment, ir: the right cell's down
// increment), along a run-time axis.
template <typename T, int RS>
__device__ __forceinline__ Flux<T> face_flux(const Params<T>& P, const Half<T>& hl, const Inc<T>& il,
                                             const Half<T>& hr, const Inc<T>& ir, int axis) {
  return face_from_sides_rt<T, RS>(P, side_of(hl, il, T(1)), side_of(hr, ir, T(-1)), axis);
}

// advance (hydro/advance.wgsl:31-52), dual_energy (hydro/dual_energy.wgsl:73-87) and the
// tenuous-gas limits (hydro/fluff.wgsl:195-205). net is the net inflow per unit dt/dx.
template <typename T>
__device__ __forceinline__ void advance(const Params<T>& P, T4<T> u1, T4<T> u2, const Flux<T>& net, T p_half, T dtdx,
                                        T4<T>& o1, T4<T>& o2) {
  T rho = tmax(u1.x + dtdx * net.m, P.rho_floor);
  T mx = u1.y + dtdx * net.mx;
  T my = u1.z + dtdx * net.my;
  T mz = u1.w + dtdx * net.mz;
  T E = u2.x + dtdx * net.e;
  T pdv = tmax(p_half, P.p_floor) * net.vn;
  T eint = u2.y + dtdx * (net.ei + pdv);
  T sc = u2.z + dtdx * net.s;
  T ke = tdiv(T(0.5) * (mx * mx + my * my + mz * mz), rho);
  T eth = E - ke;
  T emin = P.p_floor * P.igm1;
  T en0, en1;
  if (eth > P.eta_dual * E && eth >= emin && eth >= T(0)) {
    en0 = E; en1 = eth;
  } else {
    T ei = eint;
    if (!(ei >= emin)) ei = emin;
    en0 = ke + ei; en1 = ei;
  }
  if (rho < P.fluff[0]) {
    bool changed = false;
    T emax = rho * P.fluff[1];
    if (en1 > emax) { en1 = emax; changed = true; }
    T pmax = rho * P.fluff[2];
    T pm = tsqrt(mx * mx + my * my + mz * mz);
    if (pm > pmax) { T f = tdiv(pmax, pm); mx *= f; my *= f; mz *= f; changed = true; }
    if (changed) en0 = tdiv(T(0.5) * (mx * mx + my * my + mz * mz), rho) + en1;
  }
  o1.x = rho; o1.y = mx; o1.z = my; o1.w = mz;
  o2.x = en0; o2.y = en1; o2.z = sc; o2.w = T(0);
}

// Net inflow (flux.wgsl:385-394) from the lower and upper face of each axis.
template <typename T>
__device__ __forceinline__ Flux<T> net_of(const Flux<T>& xl, const Flux<T>& xh, const Flux<T>& yl, const Flux<T>& yh,
                                          const Flux<T>& zl, const Flux<T>& zh) {
  Flux<T> n;
  n.m = ((xl.m - xh.m) + (yl.m - yh.m)) + (zl.m - zh.m);
  n.mx = ((xl.mx - xh.mx) + (yl.mx - yh.mx)) + (zl.mx - zh.mx);
  n.my = ((xl.my - xh.my) + (yl.my - yh.my)) + (zl.my - zh.my);
  n.mz = ((xl.mz - xh.mz) + (yl.mz - yh.mz)) + (zl.mz - zh.mz);
  n.e = ((xl.e - xh.e) + (yl.e - yh.e)) + (zl.e - zh.e);
  n.ei = ((xl.ei - xh.ei) + (yl.ei - yh.ei)) + (zl.ei - zh.ei);
  n.s = ((xl.s - xh.s) + (yl.s - yh.s)) + (zl.s - zh.s);
  n.vn = ((xl.vn - xh.vn) + (yl.vn - yh.vn)) + (zl.vn - zh.vn);
  return n;
}

// The signal rate of a cell (hydro/rates.wgsl:60-62).
template <typename T>
__device__ __forceinline__ T rate_of(const Params<T>& P, const Prim<T>& w) {
  T signal = tsqrt(w.cs * w.cs);
  return tabs(w.vx) + tabs(w.vy) + tabs(w.vz) + P.ndim * signal;
}

// Non-negative floats order as their bit patterns: the reductions compare those.
__device__ __forceinline__ uint32_t order_bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned long long order_bits(double x) { return (unsigned long long)__double_as_longlong(x); }

// The clock pass (hydro/clock.wgsl:132-242) on one thread: bank the step just taken
// (Kahan-summed t), then dt = cfl dx / max rate, clamped by dt_max and the end time.
// clk = (dt, t, tc, steps, ...11 entries).
template <typename T>
__device__ __forceinline__ void clock_pass(const Params<T>& P, T* clk, T max_rate) {
  T prev = clk[0];
  if (prev > T(0)) {
    T y = prev - clk[2];
    T t = clk[1] + y;
    T d = t - clk[1];
    clk[2] = d - y;
    clk[1] = t;
    clk[3] = clk[3] + T(1);
  }
  T dt = P.cfl * P.dx / tmax(max_rate, T(1e-30));
  dt = tmin(dt, P.dt_max);
  T remaining = P.t_end - clk[1];
  if (remaining <= T(1e-6) * tabs(P.t_end)) dt = T(0);
  else dt = tmin(dt, remaining);
  if (T(1e30) - clk[1] < T(0.5) * dt) dt = T(0);
  if (clk[10] != T(0)) dt = T(0);
  clk[0] = dt;
}

}  // namespace shx
