/*
 * CPU restatements of Shaderous's uniform-grid Godunov hydro update. TEST INFRASTRUCTURE ONLY:
 * these functions are the parity checker for the HIP kernels in shaderous_amd/csrc. Only the
 * test suite, the smoke entry point and the benchmark's CPU-baseline leg may load them; the
 * product path never links or calls anything here.
 *
 * Two restatements live behind this header:
 *
 *  - hydro_wgsl.c: the WGSL GPU scheme (src/hydro/ WGSL files, src/eos/ideal.wgsl) with gravity,
 *    solids, species, viscosity, reference states and every other optional physics module off
 *    (the *_off stubs the hot path compiles in). Compiled twice, once with `real` = float
 *    (prefix orc_f32_, the FP32 model of the GPU path) and once with `real` = double (prefix
 *    orc_f64_, its all-FP64 twin). Every function cites the WGSL lines it follows.
 *
 *  - refhydro.c: the float64 CPU scheme RefHydro (src/reference/hydro.ts), the reference's own
 *    CPU path and the timed CPU baseline (prefix orc_ref_).
 *
 *  - riemann_exact.c: Toro's exact Riemann solver (src/reference/riemann_exact.ts).
 *
 * State layout, as the reference's readState/setState (src/sim.ts:3754-3757,3790-3807):
 *   u1[((k*ny + j)*nx + i)*4 + c] = (rho, rho vx, rho vy, rho vz)
 *   u2[((k*ny + j)*nx + i)*4 + c] = (E, rho e, rho s, mark)
 *
 * Parity status of each function is listed in the header comment of its source file.
 */
#ifndef SHX_ORACLE_H
#define SHX_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limiter and Riemann-solver codes (src/problems/types.ts:81-84; src/sim.ts hydroSlots). */
enum { ORC_LIM_KOREN = 0, ORC_LIM_MC = 1, ORC_LIM_VANLEER = 2, ORC_LIM_MINMOD = 3 };
enum { ORC_RS_HLLC = 0, ORC_RS_HLL = 1 };
/* Boundary codes (src/hydro/boundary.wgsl:18-23). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1, ORC_BC_REFLECT = 2, ORC_BC_OUTFLOW_ONLY = 3 };

/* The subset of the Params block (src/hydro/common.wgsl:4-79, packed by src/sim.ts:1720-1805)
 * that the hydro path reads with the optional physics off. Values are given in double; the
 * FP32 restatement rounds each to float once, as packParams writes a Float32Array. */
typedef struct {
  int n[3];          /* cells along x, y, z (z = 1 in 2D) */
  int bc_lo[3];
  int bc_hi[3];
  int ndim;          /* P.geom.y: the rate is sum|v| + ndim * c_s */
  int limiter;
  int riemann;
  int isentropic;    /* P.flags bit 7 (SimConfig.isentropicRecon, ideal gas) */
  double gamma;      /* P.eos.x */
  double dx;
  double cfl;
  double rho_floor;
  double p_floor;
  double eta_dual;
  double shock_thresh;
  double t_end;      /* min(tEnd, tStop), 1e30 for none */
  double dt_max;     /* 1e30 */
  double fluff[3];   /* P.fluff.xyz: (density below which the limits apply, max e, max v); 0 = off */
} orc_params;

/* ---- WGSL restatement, FP32 (orc_f32_) and FP64 (orc_f64_) ---- */
#define ORC_DECLARE(T, P)                                                                        \
  /* one full step: rates -> clock -> predict -> update. clk = (dt, t, tc, steps) as in the      \
     Clock struct; u1/u2 are replaced by the new state. work holds 24 reals per cell. */         \
  void P##step(const orc_params* prm, T* u1, T* u2, T clk[4], T* work);                          \
  /* the largest signal rate (src/hydro/rates.wgsl:33-95) */                                     \
  T P##max_rate(const orc_params* prm, const T* u1, const T* u2);                                \
  /* the clock pass (src/hydro/clock.wgsl:132-242) given the reduced rate */                     \
  void P##clock(const orc_params* prm, T clk[4], T max_rate);                                    \
  /* the predict pass into pa/pb (4 reals per cell each) using clk[0] as dt */                   \
  void P##predict(const orc_params* prm, const T* u1, const T* u2, T dt, T* pa, T* pb);          \
  /* the update pass: new state into o1/o2 */                                                    \
  void P##update(const orc_params* prm, const T* u1, const T* u2, const T* pa, const T* pb,      \
                 T dt, T* o1, T* o2);                                                            \
  /* face mass fluxes as the DEBUG_FLUX override writes them (src/hydro/update.wgsl:313-317):   \
     per cell (x lo, x hi, y lo, y hi, z lo, z hi) */                                            \
  void P##face_mass(const orc_params* prm, const T* u1, const T* u2, const T* pa, const T* pb,   \
                    T* out6);                                                                    \
  /* unit-level entry points */                                                                  \
  void P##edge4(int limiter, const T far[4], const T near[4], const T scale[4], T out[4]);       \
  void P##prim(const orc_params* prm, const T u1[4], const T u2[4], T a[4], T b[4]);            \
  /* Riemann flux between two face states given as (rho, vn, vt1, vt2, p, s); out =            \
     (m, mom n, mom t1, mom t2, e, ei, s, vn) in the face frame */                               \
  void P##riemann(const orc_params* prm, const T l[6], const T r[6], T w_hll, T out[8]);         \
  /* the Hancock predictor on a 7-cell stencil of primitives (a=rho,v; b=p,s,cs,solid) given as \
     q, xm, xp, ym, yp, zm, zp, each 8 reals; out = (a, b) 8 reals */                            \
  void P##predictor(const orc_params* prm, const T* stencil56, T dt, T out[8]);

/* ---- block AMR (src/amr): the block table (amr_tree.c) ---- */
#define ORC_BT_STRIDE 32   /* ints per block, as src/amr/tree.ts:42-53 */
#define ORC_BS 16          /* cells per block side */
#define ORC_PAD 20         /* stored texels per block side: 2 ghost layers each way */
typedef struct {
  int nblocks;
  int root[3];
  int bc_lo[3], bc_hi[3];
  int* blk;   /* per block (level, ix, iy, iz) */
  int* tab;   /* per block ORC_BT_STRIDE ints: [0] level [1..3] coords [4] flags (1 active, 2 leaf)
                 [5] parent [6..11] face neighbours [12..19] children [20..25] coarser leaves */
  int ndim;   /* 3, or 2: blocks of 16^2 cells (20^2 texels), 4 children, z faces the block itself */
} orc_amr;
/* 0, or -1 - s where block s breaks 2:1 balance */
int orc_amr_build(orc_amr* t, const int* blocks, int nblocks, const int root[3], const int bc_lo[3], const int bc_hi[3]);
int orc_amr_build_nd(orc_amr* t, const int* blocks, int nblocks, const int root[3], const int bc_lo[3],
                     const int bc_hi[3], int ndim);
void orc_amr_free(orc_amr* t);
const int* orc_amr_table(const orc_amr* t);

#define ORC_AMR_DECLARE(T, P)                                                                    \
  /* one AMR step (assemble -> clock -> predict -> update, src/amr/amr_sim.ts:2850-2929):      \
     u1/u2 hold 16^3 cells of 4 values per block (x fastest), in place; dx0 is the level-0     \
     cell size; compensated != 0 sums the fine fluxes of a coarse face with TwoSum. */         \
  void P##amr_step(const orc_params* prm, const orc_amr* t, double dx0, int compensated, T* u1, \
                   T* u2, T clk[4]);                                                           \
  /* the assemble pass alone: W (nblocks * 20^3 texels of 8 reals: rho, vx, vy, vz, p, cs, s, 0;  \
     edges and corners left as given); returns the largest leaf rate / dx */                    \
  T P##amr_assemble(const orc_params* prm, const orc_amr* t, double dx0, const T* u1, const T* u2, \
                    T* W);                                                                      \
  /* the refinement flag of every leaf (src/amr/flag.wgsl:66-155, gas cells): Loehner's second- \
     derivative estimator of density and pressure on the assembled atlas, the largest over the   \
     leaf's cells (at least 1e-30); 0 for parents. out holds nblocks reals */                     \
  void P##amr_flag(const orc_params* prm, const orc_amr* t, double dx0, const T* u1, const T* u2,  \
                   T* out);                                                                      \
  /* prolongation of one coarse cell (src/amr/cons.wgsl:76-124) from its 7-cell stencil of    \
     conserved states (centre, -x, +x, -y, +y, -z, +z; 8 reals each; ok[k] = 0 drops it),     \
     to the fine child of parity par[3] */                                                     \
  void P##amr_prolong(const orc_params* prm, const T* cons56, const int ok[7], const int par[3], \
                      T out[8]);                                                                \
  /* re-meshing (src/amr/regrid.wgsl:40-139): per new slot b, kind[b] 0 none, 1 copy of old     \
     slot src[b], 2 prolongation of old leaf src[b] into its child oct[b], 3 restriction of old \
     parent src[b]; o1/o2 hold nnew blocks */                                                   \
  void P##amr_regrid(const orc_params* prm, const orc_amr* told, const T* u1, const T* u2, int nnew, \
                     const int* kind, const int* src, const int* oct, T* o1, T* o2);                    \
  /* per-block timesteps (src/amr/bins_*.wgsl): the bins of a cycle from each block's largest rate \
     / dx (brate) and the cycle's dt, at most kmax */                                              \
  void P##amr_bins(const orc_params* prm, const orc_amr* t, int kmax, const T* brate, T dt, int* out); \
  /* one cycle of 2^kmax substeps (src/amr/amr_sim.ts:1509-1660); given (may be NULL) fixes the   \
     bins; bins_out (may be NULL) takes them; work = (block updates, leaves << kmax) */          \
  void P##amr_cycle(const orc_params* prm, const orc_amr* t, double dx0, int compensated, int kmax,  \
                    const int* given, T* u1, T* u2, T clk[4], int* bins_out, long long work[2]);

ORC_DECLARE(float, orc_f32_)
ORC_DECLARE(double, orc_f64_)
ORC_AMR_DECLARE(float, orc_f32_)
ORC_AMR_DECLARE(double, orc_f64_)

/* ---- RefHydro restatement (float64) ---- */
typedef struct orc_ref orc_ref;
/* bc codes 0 periodic, 1 outflow, other reflecting (hydro.ts:216-220) */
orc_ref* orc_ref_create(const orc_params* prm);
void orc_ref_destroy(orc_ref* h);
/* U as 7 arrays of ncell doubles: rho, mx, my, mz, E, e_int, rho*s (hydro.ts:175) */
double* orc_ref_U(orc_ref* h, int comp);
double orc_ref_t(const orc_ref* h);
int orc_ref_steps(const orc_ref* h);
/* one step (hydro.ts:378-452); returns dt. t_end clamps as step(tEnd). */
double orc_ref_step(orc_ref* h, double t_end);
/* one step over only the cells with k in [k0, k1): the bounded CPU-baseline sample. Returns dt.
   The state is not advanced. */
double orc_ref_step_slab(orc_ref* h, int k0, int k1);
void orc_ref_set_threads(int n);

/* ---- exact Riemann solver (riemann_exact.ts:9-71) ---- */
/* sample the self-similar solution at xi = x/t; in/out = (rho, u, p) */
void orc_exact_riemann(const double L[3], const double R[3], double gamma, double xi, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
