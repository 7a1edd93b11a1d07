/*
 * TEST INFRASTRUCTURE (the oracle) and the timed CPU baseline: a plain-C restatement of RefHydro,
 * the reference's float64 CPU hydro scheme (src/reference/hydro.ts:172-458), with the ideal-gas
 * equation of state (src/eos/eos.ts:395-405). Never linked into the product.
 *
 * Semantics kept as the TypeScript has them, including the documented deltas from the WGSL GPU
 * path: the signal rate uses 3 c_s whatever the dimension (hydro.ts:254), faces have no lower
 * bound from the two cells (hydro.ts:360-375), HLLC has no vacuum fallback and forms the star
 * state as f = q / (S - S*) (hydro.ts:129-143), and Koren has no early return (hydro.ts:35-45).
 *
 * Parallel form: the TS memoises each cell's prediction on first use (hydro.ts:333). Here every
 * prediction a step needs is computed first (predict-all, OpenMP over cells) and the update
 * follows (update-all); each prediction is a pure function of the state, so the result is the
 * same. Ghost predictions (hydro.ts:335-357) are formed on the fly, as in the TS.
 *
 * Parity pins (tests/test_oracle.py): conservation to 1e-12 in a periodic 16x16x4 box over 40
 * steps (reference.test.ts:93-102), Sod L1 < 0.005 at 128 cells and < 0.7 x the 64-cell error
 * (:63-78), second order on smooth advection (:80-91), positivity at a 1e5 pressure ratio
 * (:104-112), bitwise x/y/z equality (:114-133), reflecting symmetry (:135-142), and agreement
 * with the WGSL restatement in 3D where the deltas above do not act.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define NQ 6
typedef struct {
  double q[NQ];
  double su[3][NQ];
  double sd[3][NQ];
  double flag;
  int first_order;
} Pred;

struct orc_ref {
  orc_params c;
  size_t ncell;
  double* U[7];
  double* W;        /* per cell: rho, vx, vy, vz, p, s, cs */
  Pred* memo;       /* per cell, valid where have[o] */
  unsigned char* have;
  double* slab_next[7];   /* the slab update's output, kept between calls */
  double t, dt;
  int steps;
};

static const double KOREN_SMOOTH = 0.01;
static const double ISEN_SMOOTH = 0.2, ISEN_ADIABAT = 0.01, ISEN_SHEAR = 0.1, ISEN_CAP = 0.25, ISEN_ROOM = 0.01;

static double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

/* LIMITERS (hydro.ts:30-46): (far, near, scale) */
static double lim_apply(int kind, double a, double b, double scale) {
  switch (kind) {
    case ORC_LIM_MC: return a * b > 0 ? sgn(a) * dmin(dmin(2 * fabs(a), 2 * fabs(b)), 0.5 * fabs(a + b)) : 0;
    case ORC_LIM_VANLEER: return a * b > 0 ? (2 * a * b) / (a + b) : 0;
    case ORC_LIM_MINMOD: return a * b > 0 ? sgn(a) * dmin(fabs(a), fabs(b)) : 0;
    default: {
      double far = a, near = b;
      double fr = (far + 2 * near) / 3;
      double lim = far * near > 0 ? sgn(near) * dmin(dmin(2 * fabs(far), 2 * fabs(near)), fabs(fr)) : 0;
      double s = KOREN_SMOOTH * scale;
      int as_is = (s * s >= 1e-30 && s * s < 1e30) || s == 0;
      double eta = as_is ? (far * far + near * near) / dmax(s * s, 1e-30)
                         : pow(far / fabs(s), 2) + pow(near / fabs(s), 2);
      double w = dmin(dmax(1.5 - eta, 0), 1);
      return lim + (fr - lim) * w;
    }
  }
}

/* isenInc (hydro.ts:65-83); m, q, p are primitive records (rho, vx, vy, vz, p, s, cs) */
static void isen_inc(int kind, const double* m, const double* q, const double* p, double p_inc, int up,
                     double gamma, double out[2]) {
  double rl = m[0], pl = m[4], rc = q[0], pc = q[4], rr = p[0], pr = p[4];
  if (!(rl > 0 && pl > 0 && rc > 0 && pc > 0 && rr > 0 && pr > 0)) { out[0] = 0; out[1] = 0; return; }
  double gi = 1 / gamma;
  double tl = rl * pow(pc / pl, gi) - rc;
  double tr = rr * pow(pc / pr, gi) - rc;
  double s = up ? 1 : -1;
  double d = up ? lim_apply(kind, -tl, tr, rc) : lim_apply(kind, tr, -tl, rc);
  double pf = pc + 0.5 * s * p_inc;
  double rn = up ? rr : rl;
  double rf = dmin(dmax((rc + 0.5 * s * d) * pow(dmax(pf, 1e-30) / pc, gi), dmin(rc, rn)), dmax(rc, rn));
#define FADE(x) dmin(dmax(2 - (x), 0), 1)
  double dvpq = hypot(hypot(p[1] - q[1], p[2] - q[2]), p[3] - q[3]);
  double dvqm = hypot(hypot(q[1] - m[1], q[2] - m[2]), q[3] - m[3]);
  double ks = FADE(dmax(fabs(pl - pc), fabs(pr - pc)) / (ISEN_SMOOTH * pc));
  double ka = FADE(dmax(fabs(tl), fabs(tr)) / (ISEN_ADIABAT * rc));
  double kv = FADE(dmax(dvpq, dvqm) / (ISEN_SHEAR * dmax(q[6], 1e-30)));
#undef FADE
  double bound = dmax(ISEN_CAP * dmin(fabs(rc - rl), fabs(rr - rc)), ISEN_ROOM * rc * ks * ka * kv);
  out[0] = 2 * s * (rf - rc);
  out[1] = bound;
}

static size_t idx(const orc_ref* h, int i, int j, int k) {
  return ((size_t)k * h->c.n[1] + j) * h->c.n[0] + i;
}
/* bcAxis (hydro.ts:216-220): codes 0 periodic, 1 outflow, anything else reflecting */
static int bc_axis(int i, int n, int lo, int hi, double* sign) {
  *sign = 1;
  if (i < 0) {
    if (lo == 0) return ((i % n) + n) % n;
    if (lo == 1) return 0;
    *sign = -1;
    return -1 - i;
  }
  if (i >= n) {
    if (hi == 0) return i % n;
    if (hi == 1) return n - 1;
    *sign = -1;
    return 2 * n - 1 - i;
  }
  return i;
}
/* loadQ (hydro.ts:223-231) */
static void load_q(const orc_ref* h, int i, int j, int k, double out[7]) {
  double fx, fy, fz;
  int x = bc_axis(i, h->c.n[0], h->c.bc_lo[0], h->c.bc_hi[0], &fx);
  int y = bc_axis(j, h->c.n[1], h->c.bc_lo[1], h->c.bc_hi[1], &fy);
  int z = bc_axis(k, h->c.n[2], h->c.bc_lo[2], h->c.bc_hi[2], &fz);
  const double* W = h->W + idx(h, x, y, z) * 7;
  out[0] = W[0]; out[1] = W[1] * fx; out[2] = W[2] * fy; out[3] = W[3] * fz;
  out[4] = W[4]; out[5] = W[5]; out[6] = W[6];
}

/* primitives (hydro.ts:233-257) over cells [o0, o1); returns the largest rate */
static double primitives(orc_ref* h) {
  const double g = h->c.gamma;
  double rate = 0;
  long nc = (long)h->ncell;
#pragma omp parallel for reduction(max : rate) schedule(static)
  for (long o = 0; o < nc; o++) {
    double rho = dmax(h->U[0][o], h->c.rho_floor);
    double eint = h->U[5][o];
    double p = (g - 1) * eint;
    if (!(p >= h->c.p_floor)) { p = h->c.p_floor; eint = p / (g - 1); }
    double cs = sqrt((g * p) / rho);
    double* W = h->W + o * 7;
    W[0] = rho; W[1] = h->U[1][o] / rho; W[2] = h->U[2][o] / rho; W[3] = h->U[3][o] / rho;
    W[4] = p; W[5] = h->U[6][o] / rho; W[6] = cs;
    double r = fabs(W[1]) + fabs(W[2]) + fabs(W[3]) + 3 * cs;
    if (r > rate) rate = r;
  }
  return rate;
}

/* the density increment moved toward the adiabat's (hydro.ts:260-265) */
static void isen(const orc_ref* h, double out[NQ], const double* m, const double* q, const double* p, int up) {
  if (!h->c.isentropic) return;
  double ib[2];
  isen_inc(h->c.limiter, m, q, p, out[4], up, h->c.gamma, ib);
  out[0] += dmin(dmax(ib[0] - out[0], -ib[1]), ib[1]);
}

/* scales (hydro.ts:268-270) */
static void scales(const double* q, double s[NQ]) {
  s[0] = q[0]; s[1] = q[6]; s[2] = q[6]; s[3] = q[6]; s[4] = q[4]; s[5] = 0;
}

/* predictRaw (hydro.ts:272-322) */
static void predict_raw(const orc_ref* h, int i, int j, int k, Pred* out) {
  const orc_params* c = &h->c;
  double q[7], nb[3][2][7];
  load_q(h, i, j, k, q);
  load_q(h, i - 1, j, k, nb[0][0]); load_q(h, i + 1, j, k, nb[0][1]);
  load_q(h, i, j - 1, k, nb[1][0]); load_q(h, i, j + 1, k, nb[1][1]);
  load_q(h, i, j, k - 1, nb[2][0]); load_q(h, i, j, k + 1, nb[2][1]);
  double sc[NQ];
  scales(q, sc);
  for (int d = 0; d < 3; d++) {
    const double* m = nb[d][0];
    const double* p = nb[d][1];
    for (int v = 0; v < NQ; v++) out->su[d][v] = lim_apply(c->limiter, q[v] - m[v], p[v] - q[v], sc[v]);
    isen(h, out->su[d], m, q, p, 1);
  }
  for (int d = 0; d < 3; d++) {
    const double* m = nb[d][0];
    const double* p = nb[d][1];
    for (int v = 0; v < NQ; v++) out->sd[d][v] = lim_apply(c->limiter, p[v] - q[v], q[v] - m[v], sc[v]);
    isen(h, out->sd[d], m, q, p, 0);
  }
  double s[3][NQ];
  for (int d = 0; d < 3; d++)
    for (int v = 0; v < NQ; v++) s[d][v] = 0.5 * (out->su[d][v] + out->sd[d][v]);
  double big4 = 0;
  for (int d = 0; d < 3; d++) { big4 = dmax(big4, fabs(out->su[d][4])); }
  for (int d = 0; d < 3; d++) { big4 = dmax(big4, fabs(out->sd[d][4])); }
  double fall = 0;
  for (int d = 0; d < 3; d++) fall = dmin(fall, 0.5 * out->su[d][0]);
  for (int d = 0; d < 3; d++) fall = dmin(fall, -0.5 * out->sd[d][0]);
  double conv = nb[0][1][1] - nb[0][0][1] + (nb[1][1][2] - nb[1][0][2]) + (nb[2][1][3] - nb[2][0][3]);
  double jump = -INFINITY;
  for (int d = 0; d < 3; d++) jump = dmax(jump, fabs(nb[d][1][4] - nb[d][0][4]) / dmin(nb[d][1][4], nb[d][0][4]));
  double flag = conv < 0 && jump > c->shock_thresh ? 1 : 0;

  double hh = (0.5 * h->dt) / c->dx;
  double rho = q[0];
  double v[3] = {q[1], q[2], q[3]};
  double cs = q[6];
  double dv = s[0][1] + s[1][2] + s[2][3];
#define ADV(n) (v[0] * s[0][n] + v[1] * s[1][n] + v[2] * s[2][n])
  double qh[NQ];
  qh[0] = rho - hh * (ADV(0) + rho * dv);
  for (int d = 0; d < 3; d++) qh[1 + d] = v[d] - hh * (ADV(1 + d) + s[d][4] / rho);
  qh[4] = q[4] - hh * (ADV(4) + rho * cs * cs * dv);
  qh[5] = q[5] - hh * ADV(5);
#undef ADV
  double rmin = qh[0] + fall;
  double pmin = qh[4] - 0.5 * big4;
  out->flag = flag;
  if (!(rmin > 0 && pmin > 0)) {
    for (int v2 = 0; v2 < NQ; v2++) out->q[v2] = q[v2];
    memset(out->su, 0, sizeof out->su);
    memset(out->sd, 0, sizeof out->sd);
    out->first_order = 1;
    return;
  }
  for (int v2 = 0; v2 < NQ; v2++) out->q[v2] = qh[v2];
  out->first_order = 0;
}

/* predict (hydro.ts:329-358) */
static void predict(const orc_ref* h, int i, int j, int k, Pred* out) {
  const orc_params* c = &h->c;
  if (i >= 0 && j >= 0 && k >= 0 && i < c->n[0] && j < c->n[1] && k < c->n[2]) {
    size_t o = idx(h, i, j, k);
    if (h->have[o]) { *out = h->memo[o]; return; }
    predict_raw(h, i, j, k, out);
    return;
  }
  double fx, fy, fz;
  int x = bc_axis(i, c->n[0], c->bc_lo[0], c->bc_hi[0], &fx);
  int y = bc_axis(j, c->n[1], c->bc_lo[1], c->bc_hi[1], &fy);
  int z = bc_axis(k, c->n[2], c->bc_lo[2], c->bc_hi[2], &fz);
  Pred src;
  predict(h, x, y, z, &src);
  for (int v = 0; v < NQ; v++) out->q[v] = src.q[v];
  out->q[1] *= fx; out->q[2] *= fy; out->q[3] *= fz;
  double w[7];
  load_q(h, i, j, k, w);
  double sc[NQ];
  scales(w, sc);
  static const int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int up = 1; up >= 0; up--) {
    for (int d = 0; d < 3; d++) {
      double* o = up ? out->su[d] : out->sd[d];
      memset(o, 0, sizeof(double) * NQ);
      if (src.first_order) continue;
      double m[7], pl[7];
      load_q(h, i - e[d][0], j - e[d][1], k - e[d][2], m);
      load_q(h, i + e[d][0], j + e[d][1], k + e[d][2], pl);
      for (int v = 0; v < NQ; v++)
        o[v] = up ? lim_apply(c->limiter, w[v] - m[v], pl[v] - w[v], sc[v]) : lim_apply(c->limiter, pl[v] - w[v], w[v] - m[v], sc[v]);
      isen(h, o, m, w, pl, up);
    }
  }
  out->flag = src.flag;
  out->first_order = src.first_order;
}

typedef struct { double rho, v[3], p, s, cs, eint; } RSide;
/* side (hydro.ts:360-375) */
static void side(const orc_ref* h, const Pred* pr, int axis, double sg, RSide* o) {
  double g = h->c.gamma;
  double f[NQ];
  const double* inc = sg > 0 ? pr->su[axis] : pr->sd[axis];
  for (int n = 0; n < NQ; n++) f[n] = pr->q[n] + 0.5 * sg * inc[n];
  o->rho = dmax(f[0], h->c.rho_floor);
  o->p = dmax(f[4], h->c.p_floor);
  double v[3] = {f[1], f[2], f[3]};
  o->v[0] = v[axis]; o->v[1] = v[(axis + 1) % 3]; o->v[2] = v[(axis + 2) % 3];
  o->s = f[5];
  o->cs = sqrt((g * o->p) / o->rho);
  o->eint = o->p / (g - 1);
}

/* riemannFlux (hydro.ts:102-159); out = (m, mom[3], e, ei, s, vn) */
static void riemann(const RSide* l, const RSide* r, double w_hll, int hll, double out[8]) {
  double sl = dmin(l->v[0] - l->cs, r->v[0] - r->cs);
  double sr = dmax(l->v[0] + l->cs, r->v[0] + r->cs);
  double el = l->eint + 0.5 * l->rho * (l->v[0] * l->v[0] + l->v[1] * l->v[1] + l->v[2] * l->v[2]);
  double er = r->eint + 0.5 * r->rho * (r->v[0] * r->v[0] + r->v[1] * r->v[1] + r->v[2] * r->v[2]);
  double ml[3], mr[3];
  for (int i = 0; i < 3; i++) { ml[i] = l->rho * l->v[i]; mr[i] = r->rho * r->v[i]; }
  double fpl[3] = {ml[0] * l->v[0] + l->p, ml[1] * l->v[0], ml[2] * l->v[0]};
  double fpr[3] = {mr[0] * r->v[0] + r->p, mr[1] * r->v[0], mr[2] * r->v[0]};
  double fel = (el + l->p) * l->v[0];
  double fer = (er + r->p) * r->v[0];
  double m, mom[3], e, vn;
  if (sl >= 0) {
    m = ml[0]; for (int i = 0; i < 3; i++) mom[i] = fpl[i]; e = fel; vn = l->v[0];
  } else if (sr <= 0) {
    m = mr[0]; for (int i = 0; i < 3; i++) mom[i] = fpr[i]; e = fer; vn = r->v[0];
  } else {
    double ql = l->rho * (sl - l->v[0]);
    double qr = r->rho * (sr - r->v[0]);
    double sm = (r->p - l->p + l->v[0] * ql - r->v[0] * qr) / (ql - qr);
    double inv = 1 / (sr - sl);
    double hm = (sr * ml[0] - sl * mr[0] + sl * sr * (r->rho - l->rho)) * inv;
    double hp[3];
    for (int i = 0; i < 3; i++) hp[i] = (sr * fpl[i] - sl * fpr[i] + sl * sr * (mr[i] - ml[i])) * inv;
    double he = (sr * fel - sl * fer + sl * sr * (er - el)) * inv;
    double cm, cp[3], ce;
    if (sm >= 0) {
      double f = ql / (sl - sm);
      double ms[3] = {f * sm, f * l->v[1], f * l->v[2]};
      double es = f * (el / l->rho + (sm - l->v[0]) * (sm + l->p / ql));
      cm = ml[0] + sl * (f - l->rho);
      for (int i = 0; i < 3; i++) cp[i] = fpl[i] + sl * (ms[i] - ml[i]);
      ce = fel + sl * (es - el);
    } else {
      double f = qr / (sr - sm);
      double ms[3] = {f * sm, f * r->v[1], f * r->v[2]};
      double es = f * (er / r->rho + (sm - r->v[0]) * (sm + r->p / qr));
      cm = mr[0] + sr * (f - r->rho);
      for (int i = 0; i < 3; i++) cp[i] = fpr[i] + sr * (ms[i] - mr[i]);
      ce = fer + sr * (es - er);
    }
    double w = hll ? 1 : w_hll;
    m = cm + (hm - cm) * w;
    for (int i = 0; i < 3; i++) mom[i] = cp[i] + (hp[i] - cp[i]) * w;
    e = ce + (he - ce) * w;
    vn = sm;
  }
  int upw = m >= 0;
  out[0] = m; out[1] = mom[0]; out[2] = mom[1]; out[3] = mom[2]; out[4] = e;
  out[5] = m * (upw ? l->eint / l->rho : r->eint / r->rho);
  out[6] = m * (upw ? l->s : r->s);
  out[7] = vn;
}

orc_ref* orc_ref_create(const orc_params* prm) {
  orc_ref* h = (orc_ref*)calloc(1, sizeof(orc_ref));
  h->c = *prm;
  h->ncell = (size_t)prm->n[0] * prm->n[1] * prm->n[2];
  for (int i = 0; i < 7; i++) h->U[i] = (double*)calloc(h->ncell, sizeof(double));
  h->W = (double*)calloc(h->ncell * 7, sizeof(double));
  h->have = (unsigned char*)calloc(h->ncell, 1);
  h->memo = NULL;
  return h;
}
void orc_ref_destroy(orc_ref* h) {
  if (!h) return;
  for (int i = 0; i < 7; i++) { free(h->U[i]); free(h->slab_next[i]); }
  free(h->W); free(h->have); free(h->memo); free(h);
}
double* orc_ref_U(orc_ref* h, int comp) { return h->U[comp]; }
double orc_ref_t(const orc_ref* h) { return h->t; }
int orc_ref_steps(const orc_ref* h) { return h->steps; }
void orc_ref_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* The update of cells with k in [k0, k1) from the predictions (hydro.ts:395-447), into next. */
static void update_cells(orc_ref* h, int k0, int k1, double* next[7]) {
  const orc_params* c = &h->c;
  const double g = c->gamma;
  const double dtdx = h->dt / c->dx;
  static const int OFF[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  int nx = c->n[0], ny = c->n[1];
  long total = (long)(k1 - k0) * ny * nx;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < total; q++) {
    int i = (int)(q % nx);
    int j = (int)((q / nx) % ny);
    int k = k0 + (int)(q / ((long)nx * ny));
    size_t o = idx(h, i, j, k);
    Pred self;
    predict(h, i, j, k, &self);
    double dm = 0, dmom[3] = {0, 0, 0}, de = 0, dei = 0, ds = 0, dvn = 0;
    for (int f = 0; f < 6; f++) {
      int axis = f >> 1;
      int upper = (f & 1) == 1;
      Pred nb;
      predict(h, i + OFF[f][0], j + OFF[f][1], k + OFF[f][2], &nb);
      RSide L, R;
      if (upper) { side(h, &self, axis, 1, &L); side(h, &nb, axis, -1, &R); }
      else { side(h, &nb, axis, 1, &L); side(h, &self, axis, -1, &R); }
      double w = dmax(self.flag, nb.flag);
      double F[8];
      riemann(&L, &R, w, c->riemann == ORC_RS_HLL, F);
      double sg = upper ? -1 : 1;
      dm += sg * F[0];
      for (int d = 0; d < 3; d++) dmom[(axis + d) % 3] += sg * F[1 + d];
      de += sg * F[4];
      dei += sg * F[5];
      ds += sg * F[6];
      dvn += sg * F[7];
    }
    double rho = dmax(h->U[0][o] + dtdx * dm, c->rho_floor);
    double mom[3];
    for (int d = 0; d < 3; d++) mom[d] = h->U[1 + d][o] + dtdx * dmom[d];
    double E = h->U[4][o] + dtdx * de;
    double eint = h->U[5][o] + dtdx * (dei + dmax(self.q[4], c->p_floor) * dvn);
    double ke = (0.5 * (mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2])) / rho;
    double eth = E - ke;
    double emin = c->p_floor / (g - 1);
    if (eth > c->eta_dual * E && eth >= emin) {
      eint = eth;
    } else {
      if (!(eint >= emin)) eint = emin;
      E = ke + eint;
    }
    next[0][o] = rho; next[1][o] = mom[0]; next[2][o] = mom[1]; next[3][o] = mom[2];
    next[4][o] = E; next[5][o] = eint;
    next[6][o] = h->U[6][o] + dtdx * ds;
  }
}

/* predict-all over the in-grid cells the update of [k0, k1) reads: planes k0-1 .. k1, mapped */
static void predict_planes(orc_ref* h, int k0, int k1) {
  const orc_params* c = &h->c;
  if (!h->memo) h->memo = (Pred*)malloc(sizeof(Pred) * h->ncell);
  memset(h->have, 0, h->ncell);
  int nz = c->n[2];
  unsigned char* plane = (unsigned char*)calloc(nz, 1);
  for (int k = k0 - 1; k <= k1; k++) {
    double s;
    int kk = bc_axis(k, nz, c->bc_lo[2], c->bc_hi[2], &s);
    plane[kk] = 1;
  }
  int nx = c->n[0], ny = c->n[1];
  long per = (long)nx * ny;
  for (int k = 0; k < nz; k++) {
    if (!plane[k]) continue;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < per; q++) {
      int i = (int)(q % nx), j = (int)(q / nx);
      size_t o = idx(h, i, j, k);
      predict_raw(h, i, j, k, &h->memo[o]);
    }
    for (long q = 0; q < per; q++) h->have[(size_t)k * per + q] = 1;
  }
  free(plane);
}

double orc_ref_step(orc_ref* h, double t_end) {
  const orc_params* c = &h->c;
  double rate = primitives(h);
  double rem = t_end - h->t;
  h->dt = dmin((c->cfl * c->dx) / rate, dmax(rem, 0));
  predict_planes(h, 0, c->n[2]);
  double* next[7];
  for (int i = 0; i < 7; i++) { next[i] = (double*)malloc(sizeof(double) * h->ncell); memcpy(next[i], h->U[i], sizeof(double) * h->ncell); }
  update_cells(h, 0, c->n[2], next);
  for (int i = 0; i < 7; i++) { free(h->U[i]); h->U[i] = next[i]; }
  h->t += h->dt;
  h->steps++;
  return h->dt;
}

/* primitives of the planes k0-2 .. k1+1 (mapped through the z boundaries) only: what the
 * update of the slab [k0, k1) reads. Returns their largest rate. */
static double primitives_slab(orc_ref* h, int k0, int k1) {
  int nz = h->c.n[2];
  long per = (long)h->c.n[0] * h->c.n[1];
  double rate = 0;
  unsigned char* plane = (unsigned char*)calloc(nz, 1);
  for (int k = k0 - 2; k <= k1 + 1; k++) {
    double sgn;
    plane[bc_axis(k, nz, h->c.bc_lo[2], h->c.bc_hi[2], &sgn)] = 1;
  }
  const double g = h->c.gamma;
  for (int k = 0; k < nz; k++) {
    if (!plane[k]) continue;
#pragma omp parallel for reduction(max : rate) schedule(static)
    for (long q = 0; q < per; q++) {
      long o = (long)k * per + q;
      double rho = dmax(h->U[0][o], h->c.rho_floor);
      double eint = h->U[5][o];
      double p = (g - 1) * eint;
      if (!(p >= h->c.p_floor)) { p = h->c.p_floor; eint = p / (g - 1); }
      double cs = sqrt((g * p) / rho);
      double* W = h->W + o * 7;
      W[0] = rho; W[1] = h->U[1][o] / rho; W[2] = h->U[2][o] / rho; W[3] = h->U[3][o] / rho;
      W[4] = p; W[5] = h->U[6][o] / rho; W[6] = cs;
      double r = fabs(W[1]) + fabs(W[2]) + fabs(W[3]) + 3 * cs;
      if (r > rate) rate = r;
    }
  }
  free(plane);
  return rate;
}

double orc_ref_step_slab(orc_ref* h, int k0, int k1) {
  const orc_params* c = &h->c;
  double rate = primitives_slab(h, k0, k1);
  h->dt = (c->cfl * c->dx) / rate;
  predict_planes(h, k0, k1);
  for (int i = 0; i < 7; i++)
    if (!h->slab_next[i]) h->slab_next[i] = (double*)malloc(sizeof(double) * h->ncell);
  update_cells(h, k0, k1, h->slab_next);
  return h->dt;
}
