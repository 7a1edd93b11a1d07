/*
 * TEST INFRASTRUCTURE (the oracle): a plain-C restatement of Shaderous's WGSL uniform-grid hydro
 * step, used only to check the HIP kernels. Compiled twice by oracle/Makefile: REAL=float with
 * PFX=orc_f32_ and REAL=double with PFX=orc_f64_. Built with -ffp-contract=off so every
 * operation rounds as written.
 *
 * Scope: the hot path with every optional module off (gravity, solids, species, viscosity,
 * reference state, cosmology, shearing box, geometry: the *_off stubs, each of which is an exact
 * no-op here, e.g. wall_step returns (across, 1, 1) with no well-balanced reconstruction,
 * src/gravity/gfield.wgsl:83-84, and add_diffusion returns its flux, src/hydro/diffusion.wgsl:22).
 *
 * The face flux is computed once per face with flux_between (src/hydro/flux.wgsl:254-347), as
 * the tiled update does (src/hydro/tile.wgsl:166-198); the gather form computes the same value
 * from the same inputs (flux.wgsl:252-253).
 *
 * Parity pins (tests/test_oracle.py): Toro known answers through the exact solver; limiter
 * invariants (tests/unit/reference.test.ts:38-59); HLLC/HLL closed forms (supersonic upwinding,
 * a stationary contact carries no mass, HLL = HLLC at w = 1, HLLC = HLL on a vacuum face);
 * conservation to round-off in periodic boxes; x/y/z permutation equality; reflecting symmetry;
 * the Sod tube against the exact solution at the reference's measured L1; and agreement with the
 * independent RefHydro restatement (refhydro.c) where the documented WGSL/RefHydro deltas do not
 * act. No function here is left unpinned.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef REAL real;
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(PFX, name)
#define RL(x) ((real)(x))
#if REAL_IS_DOUBLE
#define RFLOOR floor
#define RLOG2 log2
#else
#define RFLOOR floorf
#define RLOG2 log2f
#endif

/* Q: primitive record (src/hydro/q.wgsl:56): a = (rho, vx, vy, vz), b = (p, s, c_s, solid). */
typedef struct { real a[4]; real b[4]; } Q;

/* The Params values as the shader sees them (f32 in the FP32 build). */
typedef struct {
  int n[3], bc_lo[3], bc_hi[3], ndim, limiter, riemann, isen;
  real gamma, gm1, dx, cfl, rho_floor, p_floor, eta_dual, shock_thresh, t_end, dt_max;
  real fluff[3];
} P_t;

static P_t load_params(const orc_params* p) {
  P_t P;
  for (int a = 0; a < 3; a++) { P.n[a] = p->n[a]; P.bc_lo[a] = p->bc_lo[a]; P.bc_hi[a] = p->bc_hi[a]; }
  P.ndim = p->ndim; P.limiter = p->limiter; P.riemann = p->riemann; P.isen = p->isentropic;
  P.gamma = RL(p->gamma);
  P.gm1 = P.gamma - RL(1.0);          /* eos_gm1(): P.eos.x - 1.0 (src/eos/ideal.wgsl:230-233) */
  P.dx = RL(p->dx); P.cfl = RL(p->cfl); P.rho_floor = RL(p->rho_floor); P.p_floor = RL(p->p_floor);
  P.eta_dual = RL(p->eta_dual); P.shock_thresh = RL(p->shock_thresh);
  P.t_end = RL(p->t_end); P.dt_max = RL(p->dt_max);
  for (int a = 0; a < 3; a++) P.fluff[a] = RL(p->fluff[a]);
  return P;
}

static real rmax(real a, real b) { return a > b ? a : b; }
static real rmin(real a, real b) { return a < b ? a : b; }
static real rabs(real a) { return a < 0 ? -a : a; }
static real rsign(real a) { return a > 0 ? RL(1.0) : (a < 0 ? RL(-1.0) : RL(0.0)); }
static real rclamp(real x, real lo, real hi) { return rmin(rmax(x, lo), hi); }
#if REAL_IS_DOUBLE
static real rsqrt_(real x) { return sqrt(x); }
static real rpow(real a, real b) { return pow(a, b); }
#else
static real rsqrt_(real x) { return sqrtf(x); }
static real rpow(real a, real b) { return powf(a, b); }
#endif
/* WGSL mix(x, y, a) = x * (1 - a) + y * a */
static real rmix(real x, real y, real a) { return x * (RL(1.0) - a) + y * a; }

/* ---------------- equation of state: src/eos/ideal.wgsl:235-267 ---------------- */
static real eos_sound_speed(const P_t* P, real rho, real p) { return rsqrt_(P->gamma * p / rho); }
static real eos_eint_from_p(const P_t* P, real p) { return p / P->gm1; }

/* prim_from_cons (src/hydro/prim.wgsl:131-156) with the ideal gas and pressure_floor without a
 * Jeans floor (src/hydro/pfloor.wgsl:169-171). Returns Q as cell_q does (recon.wgsl:23-33):
 * b = (p, scalar, c_s, mark). */
static Q prim_from_cons(const P_t* P, const real* u1, const real* u2) {
  Q q;
  real rho = rmax(u1[0], P->rho_floor);
  real eint = u2[1];
  real p = P->gm1 * eint;
  real cs = eos_sound_speed(P, rho, p);
  real pmin = P->p_floor;
  if (!(p >= pmin)) {
    p = pmin;
    eint = eos_eint_from_p(P, p);
    cs = eos_sound_speed(P, rho, p);
  }
  (void)eint;
  q.a[0] = rho;
  q.a[1] = u1[1] / rho; q.a[2] = u1[2] / rho; q.a[3] = u1[3] / rho;
  q.b[0] = p; q.b[1] = u2[2] / rho; q.b[2] = cs; q.b[3] = u2[3];
  return q;
}

/* ---------------- boundary mapping: src/hydro/common.wgsl:433-539, boundary.wgsl:47-84 ------- */
typedef struct { int c[3]; real flip[3]; real side[3]; } Mapped;

static int bc_wall(int k) { return k == 2 || k == 4 || k == 5; }
static int bc_ghost_cell(int i, int n, int k) {
  int up = i >= n;
  if (k == 0) { int v = up ? i - n : i + n; return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }
  if (bc_wall(k)) return up ? 2 * n - 1 - i : -1 - i;
  return up ? n - 1 : 0;
}
static Mapped bc_map(const P_t* P, const int c0[3]) {
  Mapped m;
  for (int a = 0; a < 3; a++) { m.c[a] = c0[a]; m.flip[a] = RL(1.0); m.side[a] = RL(0.0); }
  for (int a = 0; a < 3; a++) {
    if (c0[a] >= 0 && c0[a] < P->n[a]) continue;
    int up = c0[a] >= P->n[a];
    int k = up ? P->bc_hi[a] : P->bc_lo[a];
    m.c[a] = bc_ghost_cell(c0[a], P->n[a], k);
    if (bc_wall(k)) {
      /* bc_vsign: line-tied and moving walls reverse all, reflecting the normal component */
      if (k == 4 || k == 5) { m.flip[0] = -m.flip[0]; m.flip[1] = -m.flip[1]; m.flip[2] = -m.flip[2]; }
      else m.flip[a] = -m.flip[a];
    }
    m.side[a] = (k == 3) ? (up ? RL(1.0) : RL(-1.0)) : RL(0.0);
  }
  return m;
}
/* bc_velocity (common.wgsl:536-539) without moving walls */
static void bc_velocity(const Mapped* m, const real v[3], real out[3]) {
  for (int a = 0; a < 3; a++) {
    real w = v[a] * m->flip[a] + RL(0.0);
    if (m->side[a] != RL(0.0)) w = rmax(w * m->side[a], RL(0.0)) * m->side[a];
    out[a] = w;
  }
}
static size_t cidx(const P_t* P, const int c[3]) {
  return ((size_t)c[2] * P->n[1] + c[1]) * P->n[0] + c[0];
}

/* load_q (src/hydro/recon.wgsl:37-52) */
static Q load_q(const P_t* P, const real* u1, const real* u2, const int c[3]) {
  Mapped m = bc_map(P, c);
  size_t o = cidx(P, m.c) * 4;
  Q w = prim_from_cons(P, u1 + o, u2 + o);
  Q q = w;
  bc_velocity(&m, w.a + 1, q.a + 1);
  return q;
}
/* load_pred (src/hydro/update.wgsl:47-63) */
static Q load_pred(const P_t* P, const real* pa, const real* pb, const int c[3]) {
  Mapped m = bc_map(P, c);
  size_t o = cidx(P, m.c) * 4;
  Q q;
  q.a[0] = pa[o];
  bc_velocity(&m, pa + o + 1, q.a + 1);
  for (int i = 0; i < 4; i++) q.b[i] = pb[o + i];
  return q;
}

/* ---------------- limiters: src/hydro/limiters/ *.wgsl ---------------- */
#define SMOOTH RL(0.01)
/* koren.wgsl:274-292 */
static void edge4_koren(const real far[4], const real near[4], const real scale[4], real out[4]) {
  real freev[4], s[4], s2[4], ss[4], q[4], qk[4];
  int as_is[4];
  int all_smooth = 1;
  for (int i = 0; i < 4; i++) {
    freev[i] = (far[i] + RL(2.0) * near[i]) / RL(3.0);
    s[i] = SMOOTH * scale[i];
    s2[i] = s[i] * s[i];
    ss[i] = rmax(s2[i], RL(1e-30));
    q[i] = far[i] * far[i] + near[i] * near[i];
    as_is[i] = (s2[i] >= RL(1e-30) && s2[i] < RL(1e30)) || (s[i] == RL(0.0));
    real k = RL(1.0) / rmax(rabs(s[i]), RL(1e-37));
    real fk = far[i] * k, nk = near[i] * k;
    qk[i] = fk * fk + nk * nk;
    int smooth = as_is[i] ? (q[i] <= RL(0.5) * ss[i]) : (qk[i] <= RL(0.5));
    all_smooth &= smooth;
  }
  if (all_smooth) { for (int i = 0; i < 4; i++) out[i] = freev[i]; return; }
  for (int i = 0; i < 4; i++) {
    real m = rmin(rmin(RL(2.0) * rabs(far[i]), RL(2.0) * rabs(near[i])), rabs(freev[i]));
    real lim = (rsign(far[i]) * rsign(near[i]) > RL(0.0)) ? rsign(near[i]) * m : RL(0.0);
    real eta = as_is[i] ? q[i] / ss[i] : qk[i];
    out[i] = rmix(lim, freev[i], rclamp(RL(1.5) - eta, RL(0.0), RL(1.0)));
  }
}
/* mc.wgsl:299-307 */
static real limit_mc(real a, real b) {
  real m = rmin(rmin(RL(2.0) * rabs(a), RL(2.0) * rabs(b)), RL(0.5) * rabs(a + b));
  return (rsign(a) * rsign(b) > RL(0.0)) ? rsign(a) * m : RL(0.0);
}
/* minmod.wgsl:309-316 */
static real limit_minmod(real a, real b) {
  return (rsign(a) * rsign(b) > RL(0.0)) ? rsign(a) * rmin(rabs(a), rabs(b)) : RL(0.0);
}
/* vanleer.wgsl:318-325 */
static real limit_vanleer(real a, real b) {
  return (rsign(a) * rsign(b) > RL(0.0)) ? RL(2.0) * a * (b / (a + b)) : RL(0.0);
}
static void edge4(int lim, const real far[4], const real near[4], const real scale[4], real out[4]) {
  if (lim == ORC_LIM_KOREN) { edge4_koren(far, near, scale, out); return; }
  for (int i = 0; i < 4; i++) {
    out[i] = lim == ORC_LIM_MC ? limit_mc(far[i], near[i])
           : lim == ORC_LIM_MINMOD ? limit_minmod(far[i], near[i]) : limit_vanleer(far[i], near[i]);
  }
}
static int lim_symmetric(int lim) { return lim != ORC_LIM_KOREN; }

/* ---------------- reconstruction: src/hydro/q.wgsl ---------------- */
#define ISEN_SMOOTH RL(0.2)
#define ISEN_ADIABAT RL(0.01)
#define ISEN_SHEAR RL(0.1)
#define ISEN_CAP RL(0.25)
#define ISEN_ROOM RL(0.01)
static real len3(real x, real y, real z) { return rsqrt_(x * x + y * y + z * z); }
/* isen_inc (q.wgsl:183-209) with no reference state and no potential (g_q* = 0, df = 0) */
static void isen_inc(const P_t* P, const Q* l, const Q* c, const Q* r, real bx, real sgn, real out[2]) {
  real gi = RL(1.0) / P->gamma;
  real wl0 = l->a[0], wl1 = l->b[0], wc0 = c->a[0], wc1 = c->b[0], wr0 = r->a[0], wr1 = r->b[0];
  if (!(wl0 > 0 && wl1 > 0 && wc0 > 0 && wc1 > 0 && wr0 > 0 && wr1 > 0)) { out[0] = 0; out[1] = 0; return; }
  real tl = wl0 * rpow(wc1 / wl1, gi) - wc0;
  real tr = wr0 * rpow(wc1 / wr1, gi) - wc0;
  real far = sgn > 0 ? -tl : tr;
  real near = sgn > 0 ? tr : -tl;
  real f4[4] = {far, 0, 0, 0}, n4[4] = {near, 0, 0, 0}, s4[4] = {wc0, 1, 1, 1}, d4[4];
  edge4(P->limiter, f4, n4, s4, d4);
  real d = d4[0];
  real pf = wc1 + RL(0.5) * sgn * (bx - RL(0.0) * RL(0.0));
  real wn = sgn > 0 ? wr0 : wl0;
  real rf = rclamp((wc0 + RL(0.5) * sgn * d) * rpow(rmax(pf, RL(1e-30)) / wc1, gi), rmin(wc0, wn), rmax(wc0, wn));
  real rs = RL(0.0);
  real ks = rclamp(RL(2.0) - rmax(rabs(wl1 - wc1), rabs(wr1 - wc1)) / (ISEN_SMOOTH * wc1), RL(0.0), RL(1.0));
  real ka = rclamp(RL(2.0) - rmax(rabs(tl), rabs(tr)) / (ISEN_ADIABAT * wc0), RL(0.0), RL(1.0));
  real dvr = len3(r->a[1] - c->a[1], r->a[2] - c->a[2], r->a[3] - c->a[3]);
  real dvl = len3(c->a[1] - l->a[1], c->a[2] - l->a[2], c->a[3] - l->a[3]);
  real kv = rclamp(RL(2.0) - rmax(dvr, dvl) / (ISEN_SHEAR * rmax(c->b[2], RL(1e-30))), RL(0.0), RL(1.0));
  real bound = rmax(ISEN_CAP * rmin(rabs(wc0 - wl0), rabs(wr0 - wc0)), ISEN_ROOM * wc0 * ks * ka * kv);
  out[0] = RL(2.0) * sgn * (rf - wc0) - rs;
  out[1] = bound;
}
static real isen_blend(real x, const real i[2]) { return x + rclamp(i[0] - x, -i[1], i[1]); }

/* edge_up / edge_dn (q.wgsl:214-239) with no potential (dl = dr = df = 0). up != 0: twice the
 * increment from the cell's center to its upper face; else from its lower face to the center. */
static Q edge(const P_t* P, const Q* l, const Q* c, const Q* r, int up) {
  real bl[4], br[4], al[4], ar[4];
  for (int i = 0; i < 4; i++) {
    bl[i] = c->b[i] - l->b[i]; br[i] = r->b[i] - c->b[i];      /* wb_below, wb_above */
    al[i] = c->a[i] - l->a[i]; ar[i] = r->a[i] - c->a[i];
  }
  bl[0] = bl[0] + RL(0.5) * (l->a[0] + c->a[0]) * RL(0.0);
  br[0] = br[0] + RL(0.5) * (c->a[0] + r->a[0]) * RL(0.0);
  real bscale[4] = {c->b[0] + RL(0.0), RL(0.0), c->b[2], RL(0.0)};      /* p_scale (q.wgsl:106-109) */
  real ascale[4] = {c->a[0] + RL(0.0), c->b[2], c->b[2], c->b[2]};
  Q o;
  int sym = lim_symmetric(P->limiter);
  if (up || sym) { edge4(P->limiter, bl, br, bscale, o.b); edge4(P->limiter, al, ar, ascale, o.a); }
  else { edge4(P->limiter, br, bl, bscale, o.b); edge4(P->limiter, ar, al, ascale, o.a); }
  o.b[0] = o.b[0] - c->a[0] * RL(0.0);
  if (P->isen) {
    real ii[2];
    isen_inc(P, l, c, r, o.b[0], up ? RL(1.0) : RL(-1.0), ii);
    o.a[0] = isen_blend(o.a[0], ii);
  }
  return o;
}

/* ---------------- predictor: src/hydro/predictor.wgsl:46-209 ---------------- */
/* nb = {xm, xp, ym, yp, zm, zp}; mask multiplies each axis's slopes (1 except at a block's rim
   next to a coarser block, src/amr/predict.wgsl:205-213) */
static Q predictor_masked(const P_t* P, const Q* q, const Q nb[6], real dt, const real mask[3]) {
  Q s[3];
  real big = RL(0.0), fall = RL(0.0);
  int sym = lim_symmetric(P->limiter);
  for (int a = 0; a < 3; a++) {
    const Q* m = &nb[2 * a];
    const Q* p = &nb[2 * a + 1];
    Q u = edge(P, m, q, p, 1);
    Q sa = u;
    big = rmax(big, rabs(u.b[0] + RL(0.0)));
    if (sym) {
      fall = rmin(fall, RL(-0.5) * rabs(u.a[0] + RL(0.0)));
    } else {
      Q d = edge(P, m, q, p, 0);
      for (int i = 0; i < 4; i++) { sa.a[i] = RL(0.5) * (u.a[i] + d.a[i]); sa.b[i] = RL(0.5) * (u.b[i] + d.b[i]); }
      big = rmax(big, rabs(d.b[0] + RL(0.0)));
      fall = rmin(fall, rmin(RL(0.5) * (u.a[0] + RL(0.0)), RL(-0.5) * (d.a[0] + RL(0.0))));
    }
    for (int i = 0; i < 4; i++) { s[a].a[i] = sa.a[i] * mask[a]; s[a].b[i] = sa.b[i] * mask[a]; }
  }
  const Q *xm = &nb[0], *xp = &nb[1], *ym = &nb[2], *yp = &nb[3], *zm = &nb[4], *zp = &nb[5];
  real conv = (xp->a[1] - xm->a[1]) + (yp->a[2] - ym->a[2]) + (zp->a[3] - zm->a[3]);
  real hsx = RL(0.5) * ((xm->a[0] + q->a[0]) * RL(0.0) + (q->a[0] + xp->a[0]) * RL(0.0));
  real hsy = RL(0.5) * ((ym->a[0] + q->a[0]) * RL(0.0) + (q->a[0] + yp->a[0]) * RL(0.0));
  real hsz = RL(0.5) * ((zm->a[0] + q->a[0]) * RL(0.0) + (q->a[0] + zp->a[0]) * RL(0.0));
  real pq = RL(0.0);
  real jump = rmax(rabs(xp->b[0] - xm->b[0] + hsx) / (rmin(xp->b[0], xm->b[0]) + pq),
                   rmax(rabs(yp->b[0] - ym->b[0] + hsy) / (rmin(yp->b[0], ym->b[0]) + pq),
                        rabs(zp->b[0] - zm->b[0] + hsz) / (rmin(zp->b[0], zm->b[0]) + pq)));
  real shock = (conv < RL(0.0) && jump > P->shock_thresh) ? RL(1.0) : RL(0.0);

  real h = RL(0.5) * dt / P->dx;
  real rho = q->a[0] + RL(0.0);
  real v[3] = {q->a[1], q->a[2], q->a[3]};
  real cs = q->b[2];
  real dv = s[0].a[1] + s[1].a[2] + s[2].a[3];
  real da[4], db[4];
  for (int i = 0; i < 4; i++) {
    da[i] = v[0] * s[0].a[i] + v[1] * s[1].a[i] + v[2] * s[2].a[i];
    db[i] = v[0] * s[0].b[i] + v[1] * s[1].b[i] + v[2] * s[2].b[i];
  }
  real gradp[3] = {s[0].b[0], s[1].b[0], s[2].b[0]};
  real acoustic = RL(1.0);
  Q o;
  o.a[0] = rho - h * (da[0] + rho * dv);
  for (int i = 0; i < 3; i++) o.a[1 + i] = v[i] - h * (da[1 + i] + gradp[i] / rho) + RL(0.5) * dt * RL(0.0);
  real b0 = q->b[0] - h * (db[0] + acoustic * rho * cs * cs * dv);
  real b1 = q->b[1] - h * db[1];
  real rmin_ = o.a[0] + RL(0.0) + fall;
  real pmin_ = b0 + RL(0.0) - RL(0.5) * big;
  real first = RL(0.0);
  if (!(rmin_ > RL(0.0) && (pmin_ > RL(0.0)))) {
    for (int i = 0; i < 4; i++) o.a[i] = q->a[i];
    b0 = q->b[0]; b1 = q->b[1];
    first = RL(1.0);
  }
  o.b[0] = b0; o.b[1] = b1; o.b[2] = shock; o.b[3] = first;
  return o;
}
static Q predictor(const P_t* P, const Q* q, const Q nb[6], real dt) {
  static const real one[3] = {1, 1, 1};
  return predictor_masked(P, q, nb, dt, one);
}

/* ---------------- faces: src/hydro/flux.wgsl ---------------- */
typedef struct { real rho, v[3], p, s, cs, eint, pd; } Side;
typedef struct { real m, mom[3], e, ei, s, vn; } Flux;

static void rot3(const real v[3], int axis, real o[3]) {
  if (axis == 1) { o[0] = v[1]; o[1] = v[2]; o[2] = v[0]; }        /* v.yzx */
  else if (axis == 2) { o[0] = v[2]; o[1] = v[0]; o[2] = v[1]; }   /* v.zxy */
  else { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; }
}
static void unrot3(const real v[3], int axis, real o[3]) {
  if (axis == 1) { o[0] = v[2]; o[1] = v[0]; o[2] = v[1]; }        /* v.zxy */
  else if (axis == 2) { o[0] = v[1]; o[1] = v[2]; o[2] = v[0]; }   /* v.yzx */
  else { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; }
}

/* to_side (flux.wgsl:60-85), ideal gas, no reference state */
static Side to_side(const P_t* P, const Q* pr, const Q* s, real sgn, int axis, real lo0, real lo1) {
  real hs = RL(0.5) * sgn * (RL(1.0) - pr->b[3]);
  real a[4], b[2];
  for (int i = 0; i < 4; i++) a[i] = pr->a[i] + hs * s->a[i];
  for (int i = 0; i < 2; i++) b[i] = pr->b[i] + hs * s->b[i];
  Side o;
  o.rho = rmax(a[0] + RL(0.0), rmax(lo0, P->rho_floor));
  rot3(a + 1, axis, o.v);
  o.s = b[1];
  o.p = rmax(b[0] + RL(0.0), rmax(lo1, P->p_floor));
  o.pd = o.p;
  o.cs = eos_sound_speed(P, o.rho, o.p);
  o.eint = eos_eint_from_p(P, o.p);
  return o;
}

/* riemann_flux, HLLC (riemann/hllc.wgsl:4-85) */
static Flux riemann_hllc(const Side* l, const Side* r, real w_hll) {
  real sl = rmin(l->v[0] - l->cs, r->v[0] - r->cs);
  real sr = rmax(l->v[0] + l->cs, r->v[0] + r->cs);
  real el = l->eint + RL(0.5) * l->rho * (l->v[0] * l->v[0] + l->v[1] * l->v[1] + l->v[2] * l->v[2]);
  real er = r->eint + RL(0.5) * r->rho * (r->v[0] * r->v[0] + r->v[1] * r->v[1] + r->v[2] * r->v[2]);
  real ml[3], mr[3], fpl[3], fpr[3];
  for (int i = 0; i < 3; i++) { ml[i] = l->rho * l->v[i]; mr[i] = r->rho * r->v[i]; }
  real fml = ml[0], fmr = mr[0];
  for (int i = 0; i < 3; i++) { fpl[i] = ml[i] * l->v[0] + (i == 0 ? l->pd : RL(0.0)); fpr[i] = mr[i] * r->v[0] + (i == 0 ? r->pd : RL(0.0)); }
  real fel = (el + l->p) * l->v[0];
  real fer = (er + r->p) * r->v[0];
  Flux out;
  if (sl >= RL(0.0)) {
    out.m = fml; for (int i = 0; i < 3; i++) out.mom[i] = fpl[i]; out.e = fel; out.vn = l->v[0];
  } else if (sr <= RL(0.0)) {
    out.m = fmr; for (int i = 0; i < 3; i++) out.mom[i] = fpr[i]; out.e = fer; out.vn = r->v[0];
  } else {
    real ql = l->rho * (sl - l->v[0]);
    real qr = r->rho * (sr - r->v[0]);
    real inv = RL(1.0) / (sr - sl);
    real hm = (sr * fml - sl * fmr + sl * sr * (r->rho - l->rho)) * inv;
    real hp[3];
    for (int i = 0; i < 3; i++) hp[i] = (sr * fpl[i] - sl * fpr[i] + sl * sr * (mr[i] - ml[i])) * inv;
    real he = (sr * fel - sl * fer + sl * sr * (er - el)) * inv;
    if (!(ql < RL(0.0) && qr > RL(0.0))) {
      out.m = hm; for (int i = 0; i < 3; i++) out.mom[i] = hp[i]; out.e = he;
      out.vn = RL(0.5) * (l->v[0] + r->v[0]);
      int up0 = out.m >= RL(0.0);
      out.ei = out.m * (up0 ? l->eint / l->rho : r->eint / r->rho);
      out.s = out.m * (up0 ? l->s : r->s);
      return out;
    }
    real sm = (r->pd - l->pd + l->v[0] * ql - r->v[0] * qr) / (ql - qr);
    real cm, cp[3], ce;
    if (sm >= RL(0.0)) {
      real k = (sm - l->v[0]) / (sl - sm);
      real dr = l->rho * k;
      real f = l->rho + dr;
      cm = fml + sl * dr;
      cp[0] = fpl[0] + sl * (sm * dr + l->rho * (sm - l->v[0]));
      cp[1] = fpl[1] + sl * (dr * l->v[1]);
      cp[2] = fpl[2] + sl * (dr * l->v[2]);
      ce = fel + sl * (el * k + f * (sm - l->v[0]) * (sm + l->p / ql));
    } else {
      real k = (sm - r->v[0]) / (sr - sm);
      real dr = r->rho * k;
      real f = r->rho + dr;
      cm = fmr + sr * dr;
      cp[0] = fpr[0] + sr * (sm * dr + r->rho * (sm - r->v[0]));
      cp[1] = fpr[1] + sr * (dr * r->v[1]);
      cp[2] = fpr[2] + sr * (dr * r->v[2]);
      ce = fer + sr * (er * k + f * (sm - r->v[0]) * (sm + r->p / qr));
    }
    out.m = rmix(cm, hm, w_hll);
    for (int i = 0; i < 3; i++) out.mom[i] = rmix(cp[i], hp[i], w_hll);
    out.e = rmix(ce, he, w_hll);
    out.vn = sm;
  }
  int up = out.m >= RL(0.0);
  out.ei = out.m * (up ? l->eint / l->rho : r->eint / r->rho);
  out.s = out.m * (up ? l->s : r->s);
  return out;
}

/* riemann_flux, HLL (riemann/hll.wgsl:87-121) */
static Flux riemann_hll(const Side* l, const Side* r) {
  real sl = rmin(l->v[0] - l->cs, r->v[0] - r->cs);
  real sr = rmax(l->v[0] + l->cs, r->v[0] + r->cs);
  real el = l->eint + RL(0.5) * l->rho * (l->v[0] * l->v[0] + l->v[1] * l->v[1] + l->v[2] * l->v[2]);
  real er = r->eint + RL(0.5) * r->rho * (r->v[0] * r->v[0] + r->v[1] * r->v[1] + r->v[2] * r->v[2]);
  real ml[3], mr[3], fpl[3], fpr[3];
  for (int i = 0; i < 3; i++) { ml[i] = l->rho * l->v[i]; mr[i] = r->rho * r->v[i]; }
  real fml = ml[0], fmr = mr[0];
  for (int i = 0; i < 3; i++) { fpl[i] = ml[i] * l->v[0] + (i == 0 ? l->pd : RL(0.0)); fpr[i] = mr[i] * r->v[0] + (i == 0 ? r->pd : RL(0.0)); }
  real fel = (el + l->p) * l->v[0];
  real fer = (er + r->p) * r->v[0];
  Flux out;
  if (sl >= RL(0.0)) {
    out.m = fml; for (int i = 0; i < 3; i++) out.mom[i] = fpl[i]; out.e = fel; out.vn = l->v[0];
  } else if (sr <= RL(0.0)) {
    out.m = fmr; for (int i = 0; i < 3; i++) out.mom[i] = fpr[i]; out.e = fer; out.vn = r->v[0];
  } else {
    real ql = l->rho * (sl - l->v[0]);
    real qr = r->rho * (sr - r->v[0]);
    real inv = RL(1.0) / (sr - sl);
    out.m = (sr * fml - sl * fmr + sl * sr * (r->rho - l->rho)) * inv;
    for (int i = 0; i < 3; i++) out.mom[i] = (sr * fpl[i] - sl * fpr[i] + sl * sr * (mr[i] - ml[i])) * inv;
    out.e = (sr * fel - sl * fer + sl * sr * (er - el)) * inv;
    out.vn = (r->pd - l->pd + l->v[0] * ql - r->v[0] * qr) / (ql - qr);
  }
  int up = out.m >= RL(0.0);
  out.ei = out.m * (up ? l->eint / l->rho : r->eint / r->rho);
  out.s = out.m * (up ? l->s : r->s);
  return out;
}

/* face_flux (flux.wgsl:92-125) with no potential and no reference state */
static Flux face_flux(const P_t* P, const Q* l, const Q* ls, const Q* r, const Q* rs, int axis) {
  real w = rmax(l->b[2], r->b[2]);
  real al0 = l->a[0] + RL(0.0), al1 = l->b[0] + RL(0.0);
  real ar0 = r->a[0] + RL(0.0), ar1 = r->b[0] + RL(0.0);
  real pb = rmin(al1, ar1);
  {
    real el = al1 - RL(0.5) * al0 * RL(0.0);
    real er = ar1 + RL(0.5) * ar0 * RL(0.0);
    pb = rmin(pb, rmin(el > RL(0.0) ? el : al1, er > RL(0.0) ? er : ar1));
  }
  real lo0 = RL(0.5) * rmin(al0, ar0), lo1 = RL(0.5) * pb;
  Side sl = to_side(P, l, ls, RL(1.0), axis, lo0, lo1);
  Side sr = to_side(P, r, rs, RL(-1.0), axis, lo0, lo1);
  Flux F = P->riemann == ORC_RS_HLL ? riemann_hll(&sl, &sr) : riemann_hllc(&sl, &sr, w);
  real mom[3];
  unrot3(F.mom, axis, mom);
  for (int i = 0; i < 3; i++) F.mom[i] = mom[i];
  return F;
}

/* flux_between (flux.wgsl:254-347): the face between cell a and a + e */
static Flux flux_between(const P_t* P, const real* u1, const real* u2, const real* pa, const real* pb,
                         const int a[3], int axis) {
  int ap[3] = {a[0], a[1], a[2]}, am[3] = {a[0], a[1], a[2]}, app[3] = {a[0], a[1], a[2]};
  ap[axis] += 1; am[axis] -= 1; app[axis] += 2;
  Q wl = load_q(P, u1, u2, a), wr = load_q(P, u1, u2, ap);
  Q wll = load_q(P, u1, u2, am), wrr = load_q(P, u1, u2, app);
  Q pl = load_pred(P, pa, pb, a), pr = load_pred(P, pa, pb, ap);
  Q sl = edge(P, &wll, &wl, &wr, 1);
  Q sr = edge(P, &wl, &wr, &wrr, 0);
  return face_flux(P, &pl, &sl, &pr, &sr, axis);
}

/* ---------------- update: advance.wgsl:31-52, dual_energy.wgsl:73-87, fluff.wgsl:195-205 ------ */
static void advance(const P_t* P, const real* u1, const real* u2, const real net[8], real p_half,
                    real dt, real* o1, real* o2) {
  real dtdx = dt / P->dx;
  real rho = rmax(u1[0] + dtdx * net[0], P->rho_floor);
  real mom[3];
  for (int i = 0; i < 3; i++) mom[i] = (u1[1 + i] + RL(0.0)) + dtdx * net[1 + i];
  real E = (u2[0] + RL(0.0)) + dtdx * (net[4] + RL(0.0));
  real pdv = rmax(p_half, P->p_floor) * net[7];
  real eint = u2[1] + dtdx * (net[5] + pdv);
  real sc = u2[2] + dtdx * net[6];
  real ke = RL(0.5) * (mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]) / rho;
  real en0, en1;
  {
    real eth = E - ke;
    real emin = eos_eint_from_p(P, P->p_floor);
    if (eth > P->eta_dual * E && eth >= emin && eth >= RL(0.1) * RL(0.0)) { en0 = E; en1 = eth; }
    else {
      real ei = eint;
      if (!(ei >= emin)) ei = emin;
      en0 = ke + ei; en1 = ei;
    }
  }
  real mout[3] = {mom[0], mom[1], mom[2]};
  if (rho < P->fluff[0]) {
    int changed = 0;
    real te = en1;
    real emax = rho * P->fluff[1];
    if (te > emax) { te = emax; changed = 1; }
    real pmax = rho * P->fluff[2];
    real pm = len3(mom[0], mom[1], mom[2]);
    if (pm > pmax) { for (int i = 0; i < 3; i++) mout[i] = mom[i] * (pmax / pm); changed = 1; }
    if (changed) {
      en0 = RL(0.5) * (mout[0] * mout[0] + mout[1] * mout[1] + mout[2] * mout[2]) / rho + te;
      en1 = te;
    }
  }
  o1[0] = rho; o1[1] = mout[0]; o1[2] = mout[1]; o1[3] = mout[2];
  o2[0] = en0; o2[1] = en1; o2[2] = sc; o2[3] = RL(0.0);
}

/* ---------------- passes ---------------- */
static size_t ncell(const P_t* P) { return (size_t)P->n[0] * P->n[1] * P->n[2]; }

/* rates.wgsl:33-95 (cells marked solid are skipped; none are here) and the MAX reduction */
real FN(max_rate)(const orc_params* prm, const real* u1, const real* u2) {
  P_t P = load_params(prm);
  real best = RL(0.0);
  size_t nc = ncell(&P);
  for (size_t o = 0; o < nc; o++) {
    Q w = prim_from_cons(&P, u1 + 4 * o, u2 + 4 * o);
    real signal = rsqrt_(w.b[2] * w.b[2] * RL(1.0) * RL(1.0) + RL(0.0));
    real rate = rabs(w.a[1]) + rabs(w.a[2]) + rabs(w.a[3]) + (real)P.ndim * signal + RL(0.0);
    if (u2[4 * o + 3] <= RL(0.5) && rate > best) best = rate;
  }
  return best;
}

/* clock.wgsl:132-242 with BANK, no gravity, radiation, cosmology, rotation or masses */
void FN(clock)(const orc_params* prm, real clk[4], real max_rate) {
  P_t P = load_params(prm);
  real prev = clk[0];
  if (prev > RL(0.0)) {
    /* kahan1 (common.wgsl:366-370) */
    real y = prev - clk[2];
    real t = clk[1] + y;
    clk[2] = (t - clk[1]) - y;
    clk[1] = t;
    clk[3] = clk[3] + RL(1.0);
  }
  real dt = P.cfl * P.dx / rmax(max_rate, RL(1.0e-30));
  dt = rmin(dt, P.dt_max);
  real remaining = P.t_end - clk[1];
  if (remaining <= RL(1.0e-6) * rabs(P.t_end)) dt = RL(0.0);
  else dt = rmin(dt, remaining / RL(1.0));
  if (RL(1e30) - clk[1] < RL(0.5) * RL(1.0) * dt) dt = RL(0.0);
  clk[0] = dt;
}

static void neighbors(const P_t* P, const real* u1, const real* u2, const int c[3], Q* q, Q nb[6]) {
  *q = load_q(P, u1, u2, c);
  int flat = P->n[2] == 1;
  for (int a = 0; a < 3; a++) {
    if (a == 2 && flat) { nb[4] = *q; nb[5] = *q; continue; }   /* predict.wgsl:53-55 */
    int m[3] = {c[0], c[1], c[2]}, p[3] = {c[0], c[1], c[2]};
    m[a] -= 1; p[a] += 1;
    nb[2 * a] = load_q(P, u1, u2, m);
    nb[2 * a + 1] = load_q(P, u1, u2, p);
  }
}

/* predict.wgsl:32-119 */
void FN(predict)(const orc_params* prm, const real* u1, const real* u2, real dt, real* pa, real* pb) {
  P_t P = load_params(prm);
  for (int k = 0; k < P.n[2]; k++)
    for (int j = 0; j < P.n[1]; j++)
      for (int i = 0; i < P.n[0]; i++) {
        int c[3] = {i, j, k};
        Q q;
        Q nb[6];
        neighbors(&P, u1, u2, c, &q, nb);
        Q pr = predictor(&P, &q, nb, dt);
        size_t o = cidx(&P, c) * 4;
        for (int t = 0; t < 4; t++) { pa[o + t] = pr.a[t]; pb[o + t] = pr.b[t]; }
      }
}

static void flux_pack(const Flux* F, real out[8]) {
  out[0] = F->m; out[1] = F->mom[0]; out[2] = F->mom[1]; out[3] = F->mom[2];
  out[4] = F->e; out[5] = F->ei; out[6] = F->s; out[7] = F->vn;
}

/* update.wgsl:189-405 with the tiled faces (tile.wgsl) */
void FN(update)(const orc_params* prm, const real* u1, const real* u2, const real* pa, const real* pb,
                real dt, real* o1, real* o2) {
  P_t P = load_params(prm);
  int flat = P.n[2] == 1;
  for (int k = 0; k < P.n[2]; k++)
    for (int j = 0; j < P.n[1]; j++)
      for (int i = 0; i < P.n[0]; i++) {
        int c[3] = {i, j, k};
        real lo[3][8], hi[3][8];
        for (int a = 0; a < 3; a++) {
          if (a == 2 && flat) { for (int t = 0; t < 8; t++) { lo[a][t] = 0; hi[a][t] = 0; } continue; }
          int m[3] = {i, j, k};
          m[a] -= 1;
          Flux fl = flux_between(&P, u1, u2, pa, pb, m, a);
          Flux fh = flux_between(&P, u1, u2, pa, pb, c, a);
          flux_pack(&fl, lo[a]);
          flux_pack(&fh, hi[a]);
        }
        real net[8];
        for (int t = 0; t < 8; t++)
          net[t] = ((lo[0][t] - hi[0][t]) + (lo[1][t] - hi[1][t])) + (lo[2][t] - hi[2][t]);
        size_t o = cidx(&P, c) * 4;
        advance(&P, u1 + o, u2 + o, net, pb[o], dt, o1 + o, o2 + o);
      }
}

void FN(face_mass)(const orc_params* prm, const real* u1, const real* u2, const real* pa, const real* pb,
                   real* out6) {
  P_t P = load_params(prm);
  int flat = P.n[2] == 1;
  for (int k = 0; k < P.n[2]; k++)
    for (int j = 0; j < P.n[1]; j++)
      for (int i = 0; i < P.n[0]; i++) {
        int c[3] = {i, j, k};
        size_t o = cidx(&P, c) * 6;
        for (int a = 0; a < 3; a++) {
          if (a == 2 && flat) { out6[o + 4] = 0; out6[o + 5] = 0; continue; }
          int m[3] = {i, j, k};
          m[a] -= 1;
          out6[o + 2 * a] = flux_between(&P, u1, u2, pa, pb, m, a).m;
          out6[o + 2 * a + 1] = flux_between(&P, u1, u2, pa, pb, c, a).m;
        }
      }
}

void FN(step)(const orc_params* prm, real* u1, real* u2, real clk[4], real* work) {
  P_t P = load_params(prm);
  size_t nc = ncell(&P);
  real* pa = work;
  real* pb = work + 4 * nc;
  real* o1 = work + 8 * nc;
  real* o2 = work + 12 * nc;
  real rate = FN(max_rate)(prm, u1, u2);
  FN(clock)(prm, clk, rate);
  FN(predict)(prm, u1, u2, clk[0], pa, pb);
  FN(update)(prm, u1, u2, pa, pb, clk[0], o1, o2);
  memcpy(u1, o1, 4 * nc * sizeof(real));
  memcpy(u2, o2, 4 * nc * sizeof(real));
}

/* ---------------- unit-level entry points ---------------- */
void FN(edge4)(int limiter, const real far[4], const real near[4], const real scale[4], real out[4]) {
  edge4(limiter, far, near, scale, out);
}
void FN(prim)(const orc_params* prm, const real u1[4], const real u2[4], real a[4], real b[4]) {
  P_t P = load_params(prm);
  Q q = prim_from_cons(&P, u1, u2);
  for (int i = 0; i < 4; i++) { a[i] = q.a[i]; b[i] = q.b[i]; }
}
void FN(riemann)(const orc_params* prm, const real l[6], const real r[6], real w_hll, real out[8]) {
  P_t P = load_params(prm);
  Side sl, sr;
  sl.rho = l[0]; sl.v[0] = l[1]; sl.v[1] = l[2]; sl.v[2] = l[3]; sl.p = l[4]; sl.s = l[5];
  sr.rho = r[0]; sr.v[0] = r[1]; sr.v[1] = r[2]; sr.v[2] = r[3]; sr.p = r[4]; sr.s = r[5];
  sl.pd = sl.p; sr.pd = sr.p;
  sl.cs = eos_sound_speed(&P, sl.rho, sl.p); sl.eint = eos_eint_from_p(&P, sl.p);
  sr.cs = eos_sound_speed(&P, sr.rho, sr.p); sr.eint = eos_eint_from_p(&P, sr.p);
  Flux F = P.riemann == ORC_RS_HLL ? riemann_hll(&sl, &sr) : riemann_hllc(&sl, &sr, w_hll);
  flux_pack(&F, out);
}
void FN(predictor)(const orc_params* prm, const real* st, real dt, real out[8]) {
  P_t P = load_params(prm);
  Q q;
  Q nb[6];
  for (int i = 0; i < 4; i++) { q.a[i] = st[i]; q.b[i] = st[4 + i]; }
  for (int n = 0; n < 6; n++)
    for (int i = 0; i < 4; i++) { nb[n].a[i] = st[8 * (n + 1) + i]; nb[n].b[i] = st[8 * (n + 1) + 4 + i]; }
  Q o = predictor(&P, &q, nb, dt);
  for (int i = 0; i < 4; i++) { out[i] = o.a[i]; out[4 + i] = o.b[i]; }
}

/* ====================== block AMR: the src/amr WGSL passes ======================
 * Blocks of 16^3 cells with 2 ghost layers (BS = 16, NGH = 2, PADW = 20: src/amr/grid.wgsl:7-9);
 * in 2D (t->ndim 2, is2d() in the WGSL) blocks of 16^2 cells, 20^2 texels, the z index 0 alone.
 * Per block: U (interior, leaves only), W (primitives of every texel but edges and corners),
 * PA/PB (half steps of leaves' interiors and first ghost layer). Edge and corner texels of W are
 * never written and stay zero, as in a fresh atlas (src/amr/grid.wgsl:4-5). One global dt. */
#define BS_ ORC_BS
#define PD_ ORC_PAD
#define NG_ 2

typedef struct { real a[4]; real b[4]; } Cons;

static int bt_(const orc_amr* t, int s, int k) { return t->tab[ORC_BT_STRIDE * s + k]; }
static int leaf_(const orc_amr* t, int s) { return (bt_(t, s, 4) & 3) == 3; }
static int flat_(const orc_amr* t) { return t->ndim == 2; }
/* texels and cells per block, and the z range of a loop reaching r layers beyond the cells
   (0 alone in 2D: atl() keeps the single layer, amr/grid.wgsl:164-170) */
static size_t ntex_(const orc_amr* t) { return flat_(t) ? (size_t)PD_ * PD_ : (size_t)PD_ * PD_ * PD_; }
static size_t ncel_(const orc_amr* t) { return flat_(t) ? (size_t)BS_ * BS_ : (size_t)BS_ * BS_ * BS_; }
static int zlo_(const orc_amr* t, int r) { return flat_(t) ? 0 : -r; }
static int zhi_(const orc_amr* t, int r) { return flat_(t) ? 1 : BS_ + r; }
static size_t pidx(const orc_amr* t, int s, const int l[3]) {
  return (size_t)s * ntex_(t) + ((size_t)(flat_(t) ? 0 : l[2] + NG_) * PD_ + (l[1] + NG_)) * PD_ + (l[0] + NG_);
}
static size_t iidx(const orc_amr* t, int s, const int l[3]) {
  return (size_t)s * ncel_(t) + ((size_t)(flat_(t) ? 0 : l[2]) * BS_ + l[1]) * BS_ + l[0];
}

/* Per-block timesteps (src/amr/bins_on.wgsl): the cycle's bins and the substep being taken; the
   persistent half-step atlases PA/PB hold each block's prediction from the start of its step.
   A NULL Bins (or k NULL) is a single timestep (bins_off.wgsl). */
typedef struct {
  const int* k;        /* per slot bin */
  int sub;             /* substep of the cycle */
  const real* PA;      /* mid-step predictions, 4 reals per padded texel */
  const real* PB;
} Bins;
static int bins_on_(const Bins* B) { return B && B->k; }
static int bin_(const Bins* B, int s) { return bins_on_(B) ? B->k[s] : 0; }
/* blk_active / blk_half (bins_on.wgsl:14-23) */
static int active_(const Bins* B, int s) { return !bins_on_(B) || (B->sub & ((1 << B->k[s]) - 1)) == 0; }
static int half_(const Bins* B, int s) {
  int k = bin_(B, s);
  return k > 0 && (B->sub & ((1 << k) - 1)) == (1 << (k - 1));
}

/* ghost_src (src/amr/grid.wgsl:184-222) */
typedef struct { int kind, slot, l[3]; real flip[3], side[3]; } Src;
static Src ghost_src(const orc_amr* t, int slot, const int l[3]) {
  Src s;
  s.kind = 0; s.slot = slot;
  for (int a = 0; a < 3; a++) { s.l[a] = l[a]; s.flip[a] = RL(1.0); s.side[a] = RL(0.0); }
  int count = 0, axis = 0;
  for (int a = 0; a < 3; a++) if (l[a] < 0 || l[a] >= BS_) { count++; axis = a; }
  if (count == 0) return s;
  if (count > 1) { s.kind = 2; return s; }
  int upper = l[axis] >= BS_;
  int nb = bt_(t, slot, 6 + 2 * axis + upper);
  if (nb >= 0) {
    s.slot = nb;
    s.l[axis] = l[axis] + (upper ? -BS_ : BS_);
  } else if (nb == -1) {
    s.kind = 1;
    s.slot = bt_(t, slot, 20 + 2 * axis + upper);
  } else {
    int k = -1 - nb;
    s.l[axis] = bc_ghost_cell(l[axis], BS_, k);
    if (bc_wall(k)) {
      if (k == 4 || k == 5) { s.flip[0] = -s.flip[0]; s.flip[1] = -s.flip[1]; s.flip[2] = -s.flip[2]; }
      else s.flip[axis] = -s.flip[axis];
    }
    s.side[axis] = (k == 3) ? (upper ? RL(1.0) : RL(-1.0)) : RL(0.0);
  }
  return s;
}

static Cons u_leaf(const orc_amr* t, const real* u1, const real* u2, int slot, const int l[3]) {
  Cons c;
  size_t o = iidx(t, slot, l) * 4;
  for (int i = 0; i < 4; i++) { c.a[i] = u1[o + i]; c.b[i] = u2[o + i]; }
  return c;
}
/* u_at (src/amr/cons.wgsl:19-43): a parent's cell is the mean of its children's */
static Cons u_at(const orc_amr* t, const real* u1, const real* u2, int slot, const int l[3]) {
  if (leaf_(t, slot)) return u_leaf(t, u1, u2, slot, l);
  int h[3] = {l[0] >= BS_ / 2, l[1] >= BS_ / 2, l[2] >= BS_ / 2};
  int ch = bt_(t, slot, 12 + (h[0] | (h[1] << 1) | (h[2] << 2)));
  Cons c;
  if (ch < 0 || !leaf_(t, ch)) {
    for (int i = 0; i < 4; i++) { c.a[i] = 0; c.b[i] = 0; }
    c.a[0] = 1; c.b[0] = 1; c.b[1] = 1;
    return c;
  }
  int f[3] = {2 * (l[0] - (BS_ / 2) * h[0]), 2 * (l[1] - (BS_ / 2) * h[1]), 2 * (l[2] - (BS_ / 2) * h[2])};
  for (int i = 0; i < 4; i++) { c.a[i] = 0; c.b[i] = 0; }
  /* the children along z: 2, or 1 in 2D; each child's weight kids_w() (amr/grid.wgsl:51-53) */
  int nz = flat_(t) ? 1 : 2;
  real w = flat_(t) ? RL(0.25) : RL(0.125);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < 2; j++)
      for (int i = 0; i < 2; i++) {
        int g[3] = {f[0] + i, f[1] + j, f[2] + k};
        Cons u = u_leaf(t, u1, u2, ch, g);
        for (int q = 0; q < 4; q++) { c.a[q] = c.a[q] + RL(1.0) * u.a[q]; c.b[q] = c.b[q] + RL(1.0) * u.b[q]; }
      }
  for (int q = 0; q < 4; q++) { c.a[q] = c.a[q] * w; c.b[q] = c.b[q] * w; }
  return c;
}
/* u_near (src/amr/cons.wgsl:49-55) */
static int u_near(const orc_amr* t, const real* u1, const real* u2, int slot, const int l[3], Cons* out) {
  Src s = ghost_src(t, slot, l);
  if (s.kind != 0) { for (int i = 0; i < 4; i++) { out->a[i] = 0; out->b[i] = 0; } return 0; }
  Cons u = u_at(t, u1, u2, s.slot, s.l);
  for (int i = 0; i < 3; i++) u.a[1 + i] = u.a[1 + i] * s.flip[i] + u.a[0] * RL(0.0);
  *out = u;
  return 1;
}
/* mid_block / mid_cons (src/amr/mid_on.wgsl:10-30): a leaf part way through a longer step is
   seen through its half-step prediction (its cells and first ghost layer) */
static int mid_(const orc_amr* t, const Bins* B, int s) { return bins_on_(B) && leaf_(t, s) && !active_(B, s); }
static Cons mid_cons(const orc_amr* t, const P_t* P, const Bins* B, int slot, const int l[3]) {
  size_t o = pidx(t, slot, l) * 4;
  const real* a = B->PA + o;
  const real* b = B->PB + o;
  real rho = a[0];
  real eint = eos_eint_from_p(P, b[0]);
  Cons c;
  c.a[0] = rho; c.a[1] = rho * a[1]; c.a[2] = rho * a[2]; c.a[3] = rho * a[3];
  c.b[0] = eint + RL(0.5) * rho * ((a[1] * a[1] + a[2] * a[2]) + a[3] * a[3]);
  c.b[1] = eint; c.b[2] = rho * b[1]; c.b[3] = RL(0.0);
  return c;
}
/* limit4 of the limiter slot (MC for mc and koren: koren.wgsl:294-297) */
static real limit1(int lim, real a, real b) {
  if (lim == ORC_LIM_MINMOD) return limit_minmod(a, b);
  if (lim == ORC_LIM_VANLEER) return limit_vanleer(a, b);
  return limit_mc(a, b);
}
/* mom_slope / slope_between (src/amr/cons.wgsl:65-74) */
static Cons slope_between(const P_t* P, const Cons* uc, const Cons* m, const Cons* p) {
  Cons s;
  for (int i = 0; i < 4; i++) { s.a[i] = limit1(P->limiter, uc->a[i] - m->a[i], p->a[i] - uc->a[i]); s.b[i] = limit1(P->limiter, uc->b[i] - m->b[i], p->b[i] - uc->b[i]); }
  real vc[3], vm[3], vp[3], sv[3];
  for (int i = 0; i < 3; i++) {
    vc[i] = uc->a[1 + i] / rmax(uc->a[0], P->rho_floor);
    vm[i] = m->a[1 + i] / rmax(m->a[0], P->rho_floor);
    vp[i] = p->a[1 + i] / rmax(p->a[0], P->rho_floor);
  }
  for (int i = 0; i < 3; i++) sv[i] = limit1(P->limiter, vc[i] - vm[i], vp[i] - vc[i]);
  for (int i = 0; i < 3; i++) s.a[1 + i] = vc[i] * s.a[0] + uc->a[0] * sv[i];
  return s;
}
/* prolong_slopes / prolong_child (src/amr/cons.wgsl:83-124) */
static Cons prolong(const P_t* P, const Cons v[7], const int ok[7], const int par[3]) {
  const Cons* uc = &v[0];
  Cons s[3];
  for (int a = 0; a < 3; a++) {
    for (int i = 0; i < 4; i++) { s[a].a[i] = 0; s[a].b[i] = 0; }
    if (ok[1 + 2 * a] && ok[2 + 2 * a]) s[a] = slope_between(P, uc, &v[1 + 2 * a], &v[2 + 2 * a]);
  }
  real drop[3], room[3], fit = RL(1.0);
  for (int q = 0; q < 3; q++) {
    int ai = q == 0 ? 0 : -1, bi = q == 1 ? 0 : (q == 2 ? 1 : -1);
    real d = RL(0.0);
    for (int a = 0; a < 3; a++) d = d + rabs(ai >= 0 ? s[a].a[ai] : s[a].b[bi]);
    drop[q] = RL(0.25) * d;
    room[q] = RL(0.5) * rmax(ai >= 0 ? uc->a[ai] : uc->b[bi], RL(0.0));
  }
  for (int q = 0; q < 3; q++) fit = rmin(fit, rmin(RL(1.0), room[q] / rmax(drop[q], RL(1.0e-37))));
  real off[3];
  for (int a = 0; a < 3; a++) off[a] = fit * ((real)par[a] - RL(0.5)) * RL(0.5);
  Cons o;
  for (int i = 0; i < 4; i++) {
    o.a[i] = ((uc->a[i] + off[0] * s[0].a[i]) + off[1] * s[1].a[i]) + off[2] * s[2].a[i];
    o.b[i] = ((uc->b[i] + off[0] * s[0].b[i]) + off[1] * s[1].b[i]) + off[2] * s[2].b[i];
  }
  return o;
}
/* from_coarse (src/amr/assemble.wgsl:47-75); a coarser leaf half way through its step is
   interpolated from its half-step prediction (:60-68) */
static Cons from_coarse_b(const orc_amr* t, const P_t* P, const real* u1, const real* u2, int slot, const int l[3],
                          int cs, const Bins* B);
static Cons from_coarse(const orc_amr* t, const P_t* P, const real* u1, const real* u2, int slot, const int l[3], int cs) {
  return from_coarse_b(t, P, u1, u2, slot, l, cs, NULL);
}
static Cons from_coarse_b(const orc_amr* t, const P_t* P, const real* u1, const real* u2, int slot, const int l[3],
                          int cs, const Bins* B) {
  int gf[3], lc[3], par[3];
  int Lc = bt_(t, cs, 0);
  for (int a = 0; a < 3; a++) {
    gf[a] = bt_(t, slot, 1 + a) * BS_ + l[a];
    int nc = t->root[a] * BS_ * (1 << Lc);
    int gc = gf[a] >> 1;
    gc = ((gc % nc) + nc) % nc;
    lc[a] = gc - bt_(t, cs, 1 + a) * BS_;
    par[a] = gf[a] & 1;
  }
  Cons v[7];
  int ok[7];
  int mid = mid_(t, B, cs);
  /* the centre and its neighbours along the axes there are (x and y in 2D: the z neighbours
     are left unread, not ok; assemble.wgsl:53,62-69) */
  if (flat_(t)) lc[2] = 0;
  int n = flat_(t) ? 5 : 7;
  for (int k = 0; k < 7; k++) {
    int at[3] = {lc[0], lc[1], lc[2]};
    if (k > 0) at[(k - 1) / 2] += (k & 1) == 0 ? 1 : -1;
    if (k >= n) {
      for (int i = 0; i < 4; i++) { v[k].a[i] = 0; v[k].b[i] = 0; }
      ok[k] = 0;
    } else if (mid) {
      v[k] = mid_cons(t, P, B, cs, at);
      ok[k] = 1;
    } else {
      ok[k] = u_near(t, u1, u2, cs, at, &v[k]);
    }
  }
  return prolong(P, v, ok, par);
}

/* assemble (src/amr/assemble.wgsl:77-154) into W (8 reals per texel: rho, vx, vy, vz, p, cs, s, e);
   returns the largest rate / dx over leaf interiors */
static real amr_assemble_b(const orc_amr* t, const P_t* P, double dx0, const real* u1, const real* u2, real* W,
                           const Bins* B, real* brate);
static real amr_assemble(const orc_amr* t, const P_t* P, double dx0, const real* u1, const real* u2, real* W) {
  return amr_assemble_b(t, P, dx0, u1, u2, W, NULL, NULL);
}
/* With bins, only the blocks that start a step on this substep (parents included: a parent's
   bin is its fastest child's); brate (may be NULL) takes each leaf's largest rate / dx. */
static real amr_assemble_b(const orc_amr* t, const P_t* P, double dx0, const real* u1, const real* u2, real* W,
                           const Bins* B, real* brate) {
  real best = RL(0.0);
  /* blocks are independent (each writes only its own padded block); max is exact in any order */
#pragma omp parallel for schedule(dynamic, 4) reduction(max : best)
  for (int s = 0; s < t->nblocks; s++) {
    if (!active_(B, s)) continue;
    real bbest = RL(0.0);
    real dx = (real)(dx0 / (double)(1 << bt_(t, s, 0)));
    for (int z = zlo_(t, NG_); z < zhi_(t, NG_); z++)
      for (int y = -NG_; y < BS_ + NG_; y++)
        for (int x = -NG_; x < BS_ + NG_; x++) {
          int l[3] = {x, y, z};
          Src g = ghost_src(t, s, l);
          if (g.kind == 2) continue;
          Cons u = g.kind == 0 ? (mid_(t, B, g.slot) ? mid_cons(t, P, B, g.slot, g.l) : u_at(t, u1, u2, g.slot, g.l))
                               : from_coarse_b(t, P, u1, u2, s, l, g.slot, B);
          Q w = prim_from_cons(P, u.a, u.b);
          Mapped m;
          for (int a = 0; a < 3; a++) { m.flip[a] = g.flip[a]; m.side[a] = g.side[a]; }
          real v[3];
          bc_velocity(&m, w.a + 1, v);
          real* o = W + pidx(t, s, l) * 8;
          o[0] = w.a[0]; o[1] = v[0]; o[2] = v[1]; o[3] = v[2];
          o[4] = w.b[0]; o[5] = w.b[2]; o[6] = w.b[1]; o[7] = RL(0.0);
          int inside = x >= 0 && y >= 0 && z >= 0 && x < BS_ && y < BS_ && z < (flat_(t) ? 1 : BS_);
          if (inside && leaf_(t, s)) {
            real signal = rsqrt_(w.b[2] * w.b[2] * RL(1.0) * RL(1.0) + RL(0.0));
            real rate = rabs(v[0]) + rabs(v[1]) + rabs(v[2]) + (real)P->ndim * signal + RL(0.0);
            real r = rate / dx;
            if (r > best) best = r;
            if (r > bbest) bbest = r;
          }
        }
    if (brate) brate[s] = bbest;
  }
  return best;
}

static Q load_w(const orc_amr* t, const real* W, int s, const int l[3]) {
  const real* o = W + pidx(t, s, l) * 8;
  Q q;
  q.a[0] = rabs(o[0]); q.a[1] = o[1]; q.a[2] = o[2]; q.a[3] = o[3];
  q.b[0] = o[4]; q.b[1] = o[6]; q.b[2] = o[5]; q.b[3] = o[0] < 0 ? RL(1.0) : RL(0.0);
  return q;
}
/* predict_cell (src/amr/predict.wgsl:90-181): mask multiplies each axis's slopes; in 2D the
   cell stands for its own z neighbours (:103-108) */
static Q predict_cell(const orc_amr* t, const P_t* P, const real* W, int s, const int l[3], const real mask[3],
                      real dt, real dx) {
  Q q = load_w(t, W, s, l), nb[6];
  for (int a = 0; a < 3; a++) {
    if (a == 2 && flat_(t)) { nb[4] = q; nb[5] = q; continue; }
    int m[3] = {l[0], l[1], l[2]}, p[3] = {l[0], l[1], l[2]};
    m[a] -= 1; p[a] += 1;
    nb[2 * a] = load_w(t, W, s, m);
    nb[2 * a + 1] = load_w(t, W, s, p);
  }
  P_t Pl = *P;
  Pl.dx = dx;
  return predictor_masked(&Pl, &q, nb, dt, mask);
}

/* predict (src/amr/predict.wgsl:183-221) over leaves' interiors and first ghost layer */
static void amr_predict_b(const orc_amr* t, const P_t* P, double dx0, const real* W, real dt, real* PA, real* PB,
                          const Bins* B);
static void amr_predict(const orc_amr* t, const P_t* P, double dx0, const real* W, real dt, real* PA, real* PB) {
  amr_predict_b(t, P, dx0, W, dt, PA, PB, NULL);
}
/* With bins: the leaves starting a step, each over its own step (blk_dt, amr/blk_dt.wgsl); a
   ghost whose source leaf is part way through a longer step was filled from its prediction and
   is predicted in place (predict.wgsl:193-197). */
static void amr_predict_b(const orc_amr* t, const P_t* P, double dx0, const real* W, real dt, real* PA, real* PB,
                          const Bins* B) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < t->nblocks; s++) {
    if (!leaf_(t, s) || !active_(B, s)) continue;
    real dtb = dt * (real)(1 << bin_(B, s));
    for (int z = zlo_(t, 1); z < zhi_(t, 1); z++)
      for (int y = -1; y < BS_ + 1; y++)
        for (int x = -1; x < BS_ + 1; x++) {
          int l[3] = {x, y, z};
          Src g = ghost_src(t, s, l);
          if (g.kind == 2) continue;
          int stale = bins_on_(B) && g.slot != s && leaf_(t, g.slot) && !active_(B, g.slot);
          int at_source = g.kind == 0 && !stale;
          int slot = at_source ? g.slot : s;
          int cl[3] = {at_source ? g.l[0] : l[0], at_source ? g.l[1] : l[1], at_source ? g.l[2] : l[2]};
          real mask[3] = {1, 1, 1};
          if (!at_source)
            for (int a = 0; a < 3; a++) if (l[a] == 0 || l[a] == BS_ - 1) mask[a] = 0;
          real dx = (real)(dx0 / (double)(1 << bt_(t, slot, 0)));
          Q pr = predict_cell(t, P, W, slot, cl, mask, dtb, dx);
          if (at_source) {
            Mapped m;
            for (int a = 0; a < 3; a++) { m.flip[a] = g.flip[a]; m.side[a] = g.side[a]; }
            real v[3];
            bc_velocity(&m, pr.a + 1, v);
            pr.a[1] = v[0]; pr.a[2] = v[1]; pr.a[3] = v[2];
          }
          size_t o = pidx(t, s, l) * 4;
          for (int i = 0; i < 4; i++) { PA[o + i] = pr.a[i]; PB[o + i] = pr.b[i]; }
        }
  }
}

/* flux_between on block s (src/hydro/flux.wgsl:254-347 with stored ghosts: amr/update.wgsl:50-58) */
static Flux amr_flux_between(const orc_amr* t, const P_t* P, const real* W, const real* PA, const real* PB, int s,
                             const int a[3], int axis) {
  int ap[3] = {a[0], a[1], a[2]}, am[3] = {a[0], a[1], a[2]}, app[3] = {a[0], a[1], a[2]};
  ap[axis] += 1; am[axis] -= 1; app[axis] += 2;
  Q wl = load_w(t, W, s, a), wr = load_w(t, W, s, ap), wll = load_w(t, W, s, am), wrr = load_w(t, W, s, app);
  Q pl, pr;
  size_t ol = pidx(t, s, a) * 4, orr = pidx(t, s, ap) * 4;
  for (int i = 0; i < 4; i++) { pl.a[i] = PA[ol + i]; pl.b[i] = PB[ol + i]; pr.a[i] = PA[orr + i]; pr.b[i] = PB[orr + i]; }
  Q sl = edge(P, &wll, &wl, &wr, 1);
  Q sr = edge(P, &wl, &wr, &wrr, 0);
  return face_flux(P, &pl, &sl, &pr, &sr, axis);
}

/* TwoSum: s + e = a + b exactly */
static void two_sum(real a, real b, real* s, real* e) {
  real x = a + b;
  real bv = x - a;
  real av = x - bv;
  *e = (a - av) + (b - bv);
  *s = x;
}

/* update (src/amr/update.wgsl:326-560) with match_faces (:207-290). With bins (share,
   :179-186): the leaves starting a step, over their own step, and the leaves half way through
   one that have a faster neighbour, which take only the second half of that neighbour's fluxes
   through the shared faces (a deposit: their own fluxes are zero, cells on no such face are
   copied). A face on a faster leaf of the same level takes half of that leaf's flux, as it
   computes it; a face on a refined neighbour the mean of the fine fluxes times the child's
   share (one half if faster, nothing in a deposit if in step). */
static void amr_update_b(const orc_amr* t, const P_t* P, double dx0, int comp, const real* u1, const real* u2,
                         const real* W, const real* PA, const real* PB, real dt, real* o1, real* o2, const Bins* B,
                         const char* lu);
static void amr_update(const orc_amr* t, const P_t* P, double dx0, int comp, const real* u1, const real* u2,
                       const real* W, const real* PA, const real* PB, real dt, real* o1, real* o2) {
  amr_update_b(t, P, dx0, comp, u1, u2, W, PA, PB, dt, o1, o2, NULL, NULL);
}
static void amr_update_b(const orc_amr* t, const P_t* P, double dx0, int comp, const real* u1, const real* u2,
                         const real* W, const real* PA, const real* PB, real dt, real* o1, real* o2, const Bins* B,
                         const char* lu) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int s = 0; s < t->nblocks; s++) {
    if (!leaf_(t, s)) continue;
    if (lu && !lu[s]) continue;
    real dx = (real)(dx0 / (double)(1 << bt_(t, s, 0)));
    int gbin = bin_(B, s);
    int deposit = bins_on_(B) && !active_(B, s);
    real dtb = dt * (real)(1 << gbin);
    for (int z = 0; z < zhi_(t, 0); z++)
      for (int y = 0; y < BS_; y++)
        for (int x = 0; x < BS_; x++) {
          int l[3] = {x, y, z};
          real fl[3][8], fh[3][8];
          for (int axis = 0; axis < 3; axis++) {
            /* no faces along z in 2D (update.wgsl:374,380) */
            if (deposit || (axis == 2 && flat_(t))) {
              for (int q = 0; q < 8; q++) { fl[axis][q] = RL(0.0); fh[axis][q] = RL(0.0); }
              continue;
            }
            int m[3] = {x, y, z};
            m[axis] -= 1;
            Flux a = amr_flux_between(t, P, W, PA, PB, s, m, axis);
            Flux b = amr_flux_between(t, P, W, PA, PB, s, l, axis);
            flux_pack(&a, fl[axis]);
            flux_pack(&b, fh[axis]);
          }
          int touched = 0;
          /* faces next to a refined neighbour: the mean of the fine fluxes; with bins, faces next
             to a faster leaf of the same level: half of its flux */
          for (int k = 0; k < (flat_(t) ? 4 : 6); k++) {
            int axis = k / 2, hi = k & 1;
            if (l[axis] != (hi ? BS_ - 1 : 0)) continue;
            int np = bt_(t, s, 6 + k);
            if (np < 0) continue;
            int upper = !hi;
            real* dst = hi ? fh[axis] : fl[axis];
            if (!leaf_(t, np)) {
              int h[3] = {l[0] >= BS_ / 2, l[1] >= BS_ / 2, l[2] >= BS_ / 2};
              int bits[3] = {h[0], h[1], h[2]};
              bits[axis] = upper;
              int ch = bt_(t, np, 12 + (bits[0] | (bits[1] << 1) | (bits[2] << 2)));
              real part = RL(1.0);
              if (bins_on_(B)) part = bin_(B, ch) < gbin ? RL(0.5) : (deposit ? RL(0.0) : RL(1.0));
              if (part == RL(0.0)) {
                for (int q = 0; q < 8; q++) dst[q] = RL(0.0);
                continue;
              }
              int a0[3] = {2 * (l[0] - (BS_ / 2) * h[0]), 2 * (l[1] - (BS_ / 2) * h[1]), 2 * (l[2] - (BS_ / 2) * h[2])};
              a0[axis] = upper ? BS_ - 1 : -1;
              int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
              /* fine faces along each transverse axis: 2, or 1 along z in 2D, where the faces
                 are edges, two fine ones per coarse one (update.wgsl:246-251) */
              int n1 = (flat_(t) && t1 == 2) ? 1 : 2, n2 = (flat_(t) && t2 == 2) ? 1 : 2;
              real w = RL(1.0) / (real)(n1 * n2);
              real sum[8] = {0, 0, 0, 0, 0, 0, 0, 0}, err[8] = {0, 0, 0, 0, 0, 0, 0, 0};
              for (int j = 0; j < n2; j++)
                for (int i = 0; i < n1; i++) {
                  int a[3] = {a0[0], a0[1], a0[2]};
                  a[t1] += i;
                  a[t2] += j;
                  Flux f = amr_flux_between(t, P, W, PA, PB, ch, a, axis);
                  real fv[8];
                  flux_pack(&f, fv);
                  for (int q = 0; q < 8; q++) {
                    if (comp) { real e; two_sum(sum[q], fv[q], &sum[q], &e); err[q] = err[q] + e; }
                    else sum[q] = sum[q] + fv[q];
                  }
                }
              for (int q = 0; q < 8; q++) dst[q] = (sum[q] + err[q]) * w * part;
              touched = 1;
            } else if (bins_on_(B) && np != s && bin_(B, np) < gbin) {
              int a[3] = {l[0], l[1], l[2]};
              a[axis] = upper ? BS_ - 1 : -1;
              Flux f = amr_flux_between(t, P, W, PA, PB, np, a, axis);
              real fv[8];
              flux_pack(&f, fv);
              for (int q = 0; q < 8; q++) dst[q] = fv[q] * RL(0.5);
              touched = 1;
            }
          }
          size_t o = iidx(t, s, l) * 4;
          if (deposit && !touched) {
            for (int q = 0; q < 4; q++) { o1[o + q] = u1[o + q]; o2[o + q] = u2[o + q]; }
            continue;
          }
          real net[8];
          for (int q = 0; q < 8; q++) net[q] = ((fl[0][q] - fh[0][q]) + (fl[1][q] - fh[1][q])) + (fl[2][q] - fh[2][q]);
          P_t Pl = *P;
          Pl.dx = dx;
          advance(&Pl, u1 + o, u2 + o, net, PB[pidx(t, s, l) * 4], dtb, o1 + o, o2 + o);
        }
  }
}

void FN(amr_step)(const orc_params* prm, const orc_amr* t, double dx0, int compensated, real* u1, real* u2,
                  real clk[4]) {
  P_t P = load_params(prm);
  size_t np = (size_t)t->nblocks * ntex_(t);
  size_t ni = (size_t)t->nblocks * ncel_(t);
  real* W = (real*)calloc(np * 8, sizeof(real));
  real* PA = (real*)calloc(np * 4, sizeof(real));
  real* PB = (real*)calloc(np * 4, sizeof(real));
  real* o1 = (real*)malloc(ni * 4 * sizeof(real));
  real* o2 = (real*)malloc(ni * 4 * sizeof(real));
  memcpy(o1, u1, ni * 4 * sizeof(real));
  memcpy(o2, u2, ni * 4 * sizeof(real));
  real rate = amr_assemble(t, &P, dx0, u1, u2, W);
  /* the clock with PER_DX (src/hydro/clock.wgsl:160-161): dx taken as 1, rates per unit dx */
  orc_params pc = *prm;
  pc.dx = 1.0;
  FN(clock)(&pc, clk, rate);
  amr_predict(t, &P, dx0, W, clk[0], PA, PB);
  amr_update(t, &P, dx0, compensated, u1, u2, W, PA, PB, clk[0], o1, o2);
  memcpy(u1, o1, ni * 4 * sizeof(real));
  memcpy(u2, o2, ni * 4 * sizeof(real));
  free(W); free(PA); free(PB); free(o1); free(o2);
}

real FN(amr_assemble)(const orc_params* prm, const orc_amr* t, double dx0, const real* u1, const real* u2, real* W) {
  P_t P = load_params(prm);
  return amr_assemble(t, &P, dx0, u1, u2, W);
}

/* The refinement flag of every leaf (src/amr/flag.wgsl:66-155 for gas cells: second :66-83,
   lohner :85-106, main :131-155 without the obstacle, species, burning, point-mass, density-
   threshold and Jeans terms), on the atlas the assembly leaves: per cell Loehner's
   chi = sqrt(sum_d (q+ - 2q + q-)^2 / max(sum_d (|q+ - q| + |q - q-| + EPS(|q+| + 2|q| + |q-|))^2,
   1e-30)) of density and of pressure over the block's axes; out[s] = the largest
   max(chi_rho, chi_p, GAS_MIN) over the leaf's cells, 0 for parents. */
static void amr_flag_w(const orc_amr* t, const real* W, real* out) {
  const real EPS = RL(0.01), GAS_MIN = RL(1e-30);
  const int axes = flat_(t) ? 2 : 3;
  for (int s = 0; s < t->nblocks; s++) {
    real best = RL(0.0);
    if (leaf_(t, s))
      for (int z = zlo_(t, 0); z < zhi_(t, 0); z++)
        for (int y = 0; y < BS_; y++)
          for (int x = 0; x < BS_; x++) {
            const int l[3] = {x, y, z};
            const real* c = W + pidx(t, s, l) * 8;
            const real q0[2] = {c[0], c[4]};
            real num[2] = {RL(0.0), RL(0.0)}, den[2] = {RL(0.0), RL(0.0)};
            for (int a = 0; a < axes; a++) {
              int lp[3] = {x, y, z}, lm[3] = {x, y, z};
              lp[a] += 1;
              lm[a] -= 1;
              const real* wp = W + pidx(t, s, lp) * 8;
              const real* wm = W + pidx(t, s, lm) * 8;
              for (int k = 0; k < 2; k++) {
                const real qp = wp[4 * k], qm = wm[4 * k], q = q0[k];
                const real d = rabs(qp - q) + rabs(q - qm) + EPS * (rabs(qp) + RL(2.0) * rabs(q) + rabs(qm));
                const real s2 = qp - RL(2.0) * q + qm;
                num[k] += s2 * s2;
                den[k] += d * d;
              }
            }
            const real chi = rmax(rsqrt_(num[0] / rmax(den[0], RL(1e-30))), rsqrt_(num[1] / rmax(den[1], RL(1e-30))));
            best = rmax(best, rmax(chi, GAS_MIN));
          }
    out[s] = best;
  }
}

void FN(amr_flag)(const orc_params* prm, const orc_amr* t, double dx0, const real* u1, const real* u2, real* out) {
  P_t P = load_params(prm);
  real* W = (real*)calloc((size_t)t->nblocks * ntex_(t) * 8, sizeof(real));
  amr_assemble(t, &P, dx0, u1, u2, W);
  amr_flag_w(t, W, out);
  free(W);
}

/* ---------------- per-block timesteps (src/amr/bins_*.wgsl, amr_sim.ts encodeCycle :1509-1660) ---------------- */

/* bins_tick.wgsl MODE 0: bank the substep just taken (Kahan-summed t) */
static void clock_bank(real clk[4]) {
  real prev = clk[0];
  if (prev > RL(0.0)) {
    real y = prev - clk[2];
    real t = clk[1] + y;
    clk[2] = (t - clk[1]) - y;
    clk[1] = t;
    clk[3] = clk[3] + RL(1.0);
  }
}
/* hydro/clock.wgsl with PER_DX, BANK off and CYCLE substeps: the cycle's dt, the t_end remainder
   shared among its substeps (:119-131) */
static void clock_set_cycle(const P_t* P, real clk[4], real max_rate, real cycle) {
  real dt = P->cfl * RL(1.0) / rmax(max_rate, RL(1.0e-30));
  dt = rmin(dt, P->dt_max);
  real remaining = P->t_end - clk[1];
  if (remaining <= RL(1.0e-6) * rabs(P->t_end)) dt = RL(0.0);
  else dt = rmin(dt, remaining / cycle);
  if (RL(1e30) - clk[1] < RL(0.5) * cycle * dt) dt = RL(0.0);
  clk[0] = dt;
}

/* The bins of a cycle: bins_init.wgsl (each leaf's largest power-of-two multiple of dt its own
   rate allows, at most kmax; parents and a held clock kmax), one sweep of bins_relax.wgsl with
   DILATE (every block down to its fastest face neighbour's bin), then (kmax + levels + 2) | 1
   sweeps without: a leaf at most one bin slower than a face neighbour (a refined neighbour
   stands for its fastest child) and never slower than a coarser neighbour; a parent its fastest
   child's bin. Sweeps are Jacobi (each reads the previous sweep's bins). */
void FN(amr_bins)(const orc_params* prm, const orc_amr* t, int kmax, const real* brate, real dt, int* out) {
  P_t P = load_params(prm);
  int n = t->nblocks, levels = 0;
  for (int s = 0; s < n; s++) if (bt_(t, s, 0) + 1 > levels) levels = bt_(t, s, 0) + 1;
  int* A = (int*)malloc(sizeof(int) * (size_t)n);
  int* Bk = (int*)malloc(sizeof(int) * (size_t)n);
  for (int s = 0; s < n; s++) {
    int k = kmax;
    if (leaf_(t, s) && dt > RL(0.0)) {
      real dtb = P.cfl / rmax(brate[s], RL(1.0e-30));
      real f = RFLOOR(RLOG2(dtb / dt));
      if (f < RL(0.0)) f = RL(0.0);
      if (f > (real)kmax) f = (real)kmax;
      k = (int)f;
    }
    A[s] = k;
  }
  int sweeps = (kmax + levels + 2) | 1;
  for (int it = 0; it <= sweeps; it++) {
    const int* kin = (it % 2 == 0) ? A : Bk;
    int* kout = (it % 2 == 0) ? Bk : A;
    int dilate = it == 0;
    for (int s = 0; s < n; s++) {
      int k = kin[s];
      if (!leaf_(t, s)) {
        k = kmax;
        for (int ch = 0; ch < 8; ch++) {
          int c = bt_(t, s, 12 + ch);
          if (c >= 0 && kin[c] < k) k = kin[c];
        }
      } else {
        for (int f = 0; f < 6; f++) {
          int nb = bt_(t, s, 6 + f);
          if (nb >= 0 && nb != s) {
            int v = kin[nb] + (dilate ? 0 : 1);
            if (v < k) k = v;
          } else if (nb == -1) {
            int v = kin[bt_(t, s, 20 + f)];
            if (v < k) k = v;
          }
        }
      }
      kout[s] = k;
    }
  }
  /* 1 + sweeps passes in all, an even count: the last wrote A */
  memcpy(out, A, sizeof(int) * (size_t)n);
  free(A);
  free(Bk);
}

/* faster_neighbor (bins_lists.wgsl:22-52): a face neighbour leaf, or a child of a refined one
   touching the shared face, in a faster bin */
static int faster_neighbor(const orc_amr* t, const int* k, int s) {
  for (int f = 0; f < 6; f++) {
    int nb = bt_(t, s, 6 + f);
    if (nb < 0 || nb == s) continue;
    if (leaf_(t, nb)) {
      if (k[nb] < k[s]) return 1;
      continue;
    }
    int axis = f >> 1, want = 1 - (f & 1);
    for (int ch = 0; ch < 8; ch++) {
      if (((ch >> axis) & 1) != want) continue;
      int c = bt_(t, nb, 12 + ch);
      if (c >= 0 && leaf_(t, c) && k[c] < k[s]) return 1;
    }
  }
  return 0;
}

/* One cycle of 2^kmax substeps (encodeCycle, amr_sim.ts:1509-1660, hydro passes): substep 0 banks
   the last step, assembles every block, sets the cycle's dt from the leaves' rates and the bins
   (or takes `given` bins when non-NULL), then each substep s: (s > 0: bank, assemble the blocks
   starting a step), predict the leaves starting a step, update those and the leaves owed a
   deposit, and copy the updated blocks back (the commit). bins_out (may be NULL) takes the bins;
   work[0] the block updates made (deposits included), work[1] what single steps would have made
   (leaves << kmax). */
void FN(amr_cycle)(const orc_params* prm, const orc_amr* t, double dx0, int compensated, int kmax, const int* given,
                   real* u1, real* u2, real clk[4], int* bins_out, long long work[2]) {
  P_t P = load_params(prm);
  int n = t->nblocks;
  size_t np = (size_t)n * ntex_(t);
  size_t ni = (size_t)n * ncel_(t);
  real* W = (real*)calloc(np * 8, sizeof(real));
  real* PA = (real*)calloc(np * 4, sizeof(real));
  real* PB = (real*)calloc(np * 4, sizeof(real));
  real* o1 = (real*)malloc(ni * 4 * sizeof(real));
  real* o2 = (real*)malloc(ni * 4 * sizeof(real));
  real* brate = (real*)calloc((size_t)n, sizeof(real));
  int* k = (int*)malloc(sizeof(int) * (size_t)n);
  char* lu = (char*)malloc((size_t)n);
  memcpy(o1, u1, ni * 4 * sizeof(real));
  memcpy(o2, u2, ni * 4 * sizeof(real));
  const int cyc = 1 << kmax;
  long long updates = 0, leaves = 0;
  for (int s = 0; s < n; s++) leaves += leaf_(t, s);
  clock_bank(clk);
  real rate = amr_assemble_b(t, &P, dx0, u1, u2, W, NULL, brate);
  clock_set_cycle(&P, clk, rate, (real)cyc);
  if (given) memcpy(k, given, sizeof(int) * (size_t)n);
  else FN(amr_bins)(prm, t, kmax, brate, clk[0], k);
  if (bins_out) memcpy(bins_out, k, sizeof(int) * (size_t)n);
  for (int sub = 0; sub < cyc; sub++) {
    Bins B = {k, sub, PA, PB};
    if (sub > 0) {
      clock_bank(clk);
      (void)amr_assemble_b(t, &P, dx0, u1, u2, W, &B, NULL);
    }
    amr_predict_b(t, &P, dx0, W, clk[0], PA, PB, &B);
    for (int s = 0; s < n; s++)
      lu[s] = leaf_(t, s) && (active_(&B, s) || (half_(&B, s) && faster_neighbor(t, k, s)));
    amr_update_b(t, &P, dx0, compensated, u1, u2, W, PA, PB, clk[0], o1, o2, &B, lu);
    for (int s = 0; s < n; s++) {
      if (!lu[s]) continue;
      updates++;
      size_t o = (size_t)s * ncel_(t) * 4;
      memcpy(u1 + o, o1 + o, ncel_(t) * 4 * sizeof(real));
      memcpy(u2 + o, o2 + o, ncel_(t) * 4 * sizeof(real));
    }
  }
  if (work) { work[0] = updates; work[1] = leaves << kmax; }
  free(W); free(PA); free(PB); free(o1); free(o2); free(brate); free(k); free(lu);
}

/* Re-meshing (src/amr/regrid.wgsl:40-139): every leaf of a new tree from the old tree's state.
   Per new slot: kind 0 not a leaf (left as given), 1 copy of old slot src, 2 prolongation of
   old leaf src to its child of octant oct, 3 restriction of old parent src (the mean of its 8
   leaf children). The prolongation's stencil is the old state's (u_near, as the ghost fill's
   from_coarse); its total energy is each child's own kinetic energy plus its interpolated
   internal energy plus the offset that keeps the children's mean the coarse cell's
   (regrid.wgsl:86-97). */
static real ke_of(const P_t* P, const Cons* u) {
  return RL(0.5) * ((u->a[1] * u->a[1] + u->a[2] * u->a[2]) + u->a[3] * u->a[3]) / rmax(u->a[0], P->rho_floor);
}
void FN(amr_regrid)(const orc_params* prm, const orc_amr* told, const real* u1, const real* u2, int nnew,
                    const int* kind, const int* src, const int* oct, real* o1, real* o2) {
  P_t P = load_params(prm);
  for (int b = 0; b < nnew; b++) {
    if (kind[b] == 0) continue;
    for (int z = 0; z < zhi_(told, 0); z++)
      for (int y = 0; y < BS_; y++)
        for (int x = 0; x < BS_; x++) {
          int l[3] = {x, y, z};
          Cons u;
          if (kind[b] == 1) {
            u = u_leaf(told, u1, u2, src[b], l);
          } else if (kind[b] == 3) {
            u = u_at(told, u1, u2, src[b], l);
          } else {
            int lc[3], par[3];
            for (int a = 0; a < 3; a++) {
              lc[a] = (l[a] >> 1) + (BS_ / 2) * ((oct[b] >> a) & 1);
              par[a] = l[a] & 1;
            }
            Cons v[7];
            int ok[7];
            /* 2D: no z neighbours, children along x and y alone (regrid.wgsl:58-62, 86-97) */
            int n = flat_(told) ? 5 : 7;
            for (int k = 0; k < 7; k++) {
              int at[3] = {lc[0], lc[1], lc[2]};
              if (k > 0) at[(k - 1) / 2] += (k & 1) == 0 ? 1 : -1;
              if (k >= n) {
                for (int i = 0; i < 4; i++) { v[k].a[i] = 0; v[k].b[i] = 0; }
                ok[k] = 0;
              } else {
                ok[k] = u_near(told, u1, u2, src[b], at, &v[k]);
              }
            }
            u = prolong(&P, v, ok, par);
            real ke = RL(0.0);
            for (int k = 0; k < (flat_(told) ? 1 : 2); k++)
              for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) {
                  int pc[3] = {i, j, k};
                  Cons ch = prolong(&P, v, ok, pc);
                  ke = ke + ke_of(&P, &ch);
                }
            ke = ke * (flat_(told) ? RL(0.25) : RL(0.125));
            u.b[0] = ke_of(&P, &u) + u.b[1] + ((v[0].b[0] - v[0].b[1]) - ke);
          }
          size_t o = iidx(told, b, l) * 4;
          for (int i = 0; i < 4; i++) { o1[o + i] = u.a[i]; o2[o + i] = u.b[i]; }
        }
  }
}

void FN(amr_prolong)(const orc_params* prm, const real* cons56, const int ok[7], const int par[3], real out[8]) {
  P_t P = load_params(prm);
  Cons v[7];
  for (int k = 0; k < 7; k++)
    for (int i = 0; i < 4; i++) { v[k].a[i] = cons56[8 * k + i]; v[k].b[i] = cons56[8 * k + 4 + i]; }
  Cons o = prolong(&P, v, ok, par);
  for (int i = 0; i < 4; i++) { out[i] = o.a[i]; out[4 + i] = o.b[i]; }
}
