/*
 * libshx: the MI355X (gfx950) HIP backend of Shaderous's Godunov hydro update, behind a C ABI.
 *
 * The entry points replace the hydro step interface of the reference's uniform-grid simulation
 * (`Sim`, src/sim.ts) and its block grid (`AmrSim`, src/amr/amr_sim.ts): a caller binds them
 * through an FFI (an N-API addon for the TypeScript `Simulation` API; ctypes in Python). Each
 * declaration names the reference member it stands in for.
 *
 * Conventions:
 *  - Every entry returns an int status: SHX_OK (0) or a negative SHX_E* code; no C++ exception
 *    crosses the ABI. shx_last_error() returns the calling thread's last message.
 *  - The sim owns its device memory and its HIP stream; shx_destroy frees them
 *    (Sim.destroy, src/sim.ts:3852-3876). Host buffers are copied in and out, never aliased.
 *  - State layout (Sim.readState / setState, src/sim.ts:3031-3065, 3790-3807): two arrays of
 *    4 floats per cell, cell ((k*ny + j)*nx + i):
 *      u1 = (rho, rho vx, rho vy, rho vz)      u2 = (E, rho e, rho s, mark)
 *    With precision 64 the same layout in doubles (the *_f64 variants).
 *  - Stepping never blocks: shx_step enqueues graph launches on the sim's stream. Reads
 *    synchronize that stream.
 */
#ifndef SHX_H
#define SHX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHX_OK 0
#define SHX_EINVAL -1     /* bad argument or unsupported configuration (Sim refusal(), sim.ts:839) */
#define SHX_ENOMEM -2     /* device allocation failed (the OOM refusal, sim.ts:2302-2311) */
#define SHX_EDEVICE -3    /* a HIP call failed, or no usable GPU */
#define SHX_ESTATE -4     /* non-finite clock or state (runToEnd's NaN check, sim.ts:3708) */

/* Limiter (SimConfig.limiter, src/problems/types.ts:81-84) and Riemann solver (SimConfig.riemann). */
enum { SHX_LIM_KOREN = 0, SHX_LIM_MC = 1, SHX_LIM_VANLEER = 2, SHX_LIM_MINMOD = 3 };
enum { SHX_RS_HLLC = 0, SHX_RS_HLL = 1 };
/* Boundary codes (src/hydro/boundary.wgsl:18-23): periodic, outflow, reflecting, outflow only. */
enum { SHX_BC_PERIODIC = 0, SHX_BC_OUTFLOW = 1, SHX_BC_REFLECT = 2, SHX_BC_OUTFLOW_ONLY = 3 };
/* Uniform-grid kernel path: PLAIN keeps the reference's passes (rates, clock, predict, tiled
 * update); FUSED is one pass per step (rates of the next step in the epilogue). A flat grid on
 * the FUSED path runs the 2D kernels of the ensembles (as a batch of one grid): in FP32 the row
 * march over 60-cell strips or the LDS tile kernel, in FP64 the tile kernel. */
enum { SHX_PATH_FUSED = 0, SHX_PATH_PLAIN = 1 };

/* The SimConfig fields the hydro path reads (src/sim.ts:48-264, defaults configFor :645-736). */
typedef struct {
  int n[3];            /* cells along x, y, z; n[2] = 1 for a 2D run */
  int ndim;            /* SimConfig.ndim: the rate is sum|v| + ndim c_s (P.geom.y) */
  int bc_lo[3];
  int bc_hi[3];
  int limiter;
  int riemann;
  int isentropic;      /* SimConfig.isentropicRecon (ideal gas) */
  int precision;       /* 32 (production) or 64 (the all-FP64 precision twin) */
  int path;            /* SHX_PATH_* */
  int device;          /* HIP device ordinal */
  double dx;           /* cell size */
  double gamma;        /* eosParams.gamma (ideal gas) */
  double cfl;
  double rho_floor;
  double p_floor;
  double eta_dual;
  double shock_thresh;
  double t_end;        /* 1e30 (or +inf) for none */
  double fluff[3];     /* SimConfig.fluff: (rho below which the limits apply, max e, max v); rho 0 = off */
  int tile[3];         /* fused kernel's column tile (x, y) and z chunk; 0 = autotuned default.
                        * A flat FP32 grid takes the 2D tiles (16,16), (28,12), (28,28) and the
                        * strips (60,8), (60,16) (the default), (60,32) */
  double org[3];       /* the domain's lower corner (domainLo, src/domain.ts:17-26; Params.org):
                        * cell (i, j, k) is centred at (i + 1/2) dx + org. Read by the
                        * diagnostics' centre of mass. Default -n[0] dx / 2 on every axis */
} shx_config;

typedef struct shx_sim shx_sim;

/* Defaults of configFor (src/sim.ts:645-736) for an n^ndim grid. */
void shx_config_default(shx_config* cfg, int nx, int ny, int nz);

/* Sim.create (src/sim.ts:2283-2409) / createSim (src/create.ts:58-61): allocate the state,
 * scratch and clock on the device and build the step graphs. */
int shx_create(const shx_config* cfg, shx_sim** out);
/* Sim.destroy (src/sim.ts:3852-3876) */
int shx_destroy(shx_sim* sim);

/* Sim.setState (src/sim.ts:3031-3065): ncell = nx*ny*nz cells of u1 and u2. Resets the clock
 * to (dt 0, t = time, steps = steps) as clockState (src/sim.ts:1588-1593). */
int shx_set_state(shx_sim* sim, const float* u1, const float* u2, size_t ncell, double time, int steps);
int shx_set_state_f64(shx_sim* sim, const double* u1, const double* u2, size_t ncell, double time, int steps);
/* Sim.readState (src/sim.ts:3790-3807) */
int shx_read_state(shx_sim* sim, float* u1, float* u2);
int shx_read_state_f64(shx_sim* sim, double* u1, double* u2);

/* Sim.step (src/sim.ts:3592) / encodeSteps (src/sim.ts:3187-3328, Simulation.encodeSteps
 * :485): enqueue `count` steps (rates -> clock -> predict -> update each, dt computed and applied
 * on the GPU) and return at once; *taken (may be null) = the steps enqueued, which is `count`
 * (the uniform grid waits on no refinement check). */
int shx_step(shx_sim* sim, int count, int* taken);
/* Wait for every enqueued step. */
int shx_sync(shx_sim* sim);
/* Sim.runToEnd (src/sim.ts:3689-3718): step in batches (from `batch`, never far past the steps
 * the clock says remain) until the clock holds at t_end (dt = 0, t > 0); *steps = the clock's
 * banked step count then (clock[3], counted from the state's start, as the reference returns).
 * SHX_EINVAL without a finite t_end; SHX_ESTATE on a NaN clock or when the count passes
 * max_steps before t_end (the reference throws for both). */
int shx_run_to_end(shx_sim* sim, int batch, int max_steps, int* steps);
/* The Clock struct (src/hydro/common.wgsl:84-96): (dt, t, tc, steps, a, ah, a1, hub, shear,
 * shear_c, hold), as the step about to be taken sees it. */
int shx_read_clock(shx_sim* sim, double out[11]);
/* Sim.diagnostics (src/sim.ts:1647-1673): the 24 channels of hydro/diagnostics.wgsl:13-98 in
 * the reference's order (DIAG_OPS, src/sim.ts:1808), every cell weighted by its volume
 * dv = dx^ndim (P.geom.x, sim.ts:1733), summed in FP64 in a fixed order (deterministic):
 * out[0] mass, [1..3] momentum, [4] kinetic energy, [5] internal energy as the total holds it
 * (E - KE), [6] gravitational energy (0 on this path), [7] max Mach, [8] floored cells
 * (rho <= 1.0001 rho_floor or p <= 1.0001 p_floor), [9] Jeans-unresolved cells (0),
 * [10, 11] rho min/max, [12, 13] temperature (gamma - 1) e / rho min/max, [14, 15] specific
 * scalar min/max, [16, 17] p min/max, [18, 19] speed min/max, [20..22] sum of rho dv x (x the
 * cell centre; z = 0 on a flat grid), [23] nuclear heating (0). */
int shx_diagnostics(shx_sim* sim, double out[24]);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The 24 channels of
 * shx_diagnostics over a state the caller holds in device memory (u1, u2: n[0] n[1] n[2] cells
 * of the layout above in cfg->precision; out: 24 doubles of device memory), enqueued on
 * `stream` (a hipStream_t; null for the default stream), returning at once. It allocates
 * nothing: the per-workgroup partial sums live in the library's file-scope device array of 256
 * records x 24 doubles (48 KiB per GPU, held until the process ends), which every call rewrites
 * in full before reading it. Calls that overlap on one GPU share that array and race: serialise
 * them (one stream, or events). At most 2^31 - 1 cells (SHX_EINVAL beyond). Same kernels and
 * summation order as shx_diagnostics: the results are bit-identical. */
int shx_diagnostics_noalloc(const shx_config* cfg, const void* u1, const void* u2, double* out, void* stream);
/* Sim.fluxMismatch (src/sim.ts:3724-3787) with the DEBUG_FLUX override (hydro/update.wgsl:187):
 * every face's mass flux as each of its two cells computed it; *mismatched = faces whose two
 * values differ in any bit, *faces = faces compared. */
int shx_flux_mismatch(shx_sim* sim, int64_t* mismatched, int64_t* faces);
/* Sim.profile (src/sim.ts:3541-3590): time `steps` steps pass by pass with HIP events.
 * ms[0] rates, [1] clock, [2] predict, [3] update (plain path) or ms[3] the fused pass;
 * ms[4] whole step. */
int shx_profile(shx_sim* sim, int steps, double ms[5]);
/* The name of the step's dominant kernel ("k_fused_r", "k_fused", "k_ens_rows", "k_ens_fused" or
 * "k_update"), for profiles. */
int shx_kernel_name(shx_sim* sim, char* buf, size_t len);

/* The autotuner: time the fused kernel's tile candidates on the current state (which it
 * restores) and keep the fastest; *chosen = (tx, ty, zc). */
int shx_autotune(shx_sim* sim, int reps, int chosen[3]);

/* Sim.writeParams (src/sim.ts:2680-2700): take the run-time fields of *cfg (gamma, cfl, t_end,
 * rho_floor, p_floor, eta_dual, shock_thresh, fluff, org) into the parameter block the steps
 * read, from the next enqueued step on (the step graphs are rebuilt). The grid, boundaries,
 * precision, path, device and scheme must equal the sim's (SHX_EINVAL otherwise: a new sim is
 * needed for those, as the reference recreates its textures). The clock and state are kept. */
int shx_write_params(shx_sim* sim, const shx_config* cfg);
/* Sim.buildHydro (src/sim.ts:2434-2560): select the hydro kernels for the limiter, Riemann
 * solver and reconstruction of *cfg (hydroSlots, sim.ts:1188-1232), with its run-time fields
 * as shx_write_params; the other fields must equal the sim's. The state and clock are kept. */
int shx_build_hydro(shx_sim* sim, const shx_config* cfg);

/* ---------------- asynchronous readbacks ----------------
 * The reference reads state and diagnostics back with mapAsync (Sim.readState / readDiagnostics,
 * src/sim.ts:3598-3640, 3790-3807): the copy is enqueued when the call is made, and the caller
 * waits on it later. A readback is that copy: begun on the sim's stream (after every step
 * enqueued so far), into host memory the readback owns, so the sim may step, re-mesh or be
 * destroyed meanwhile. `what`: bit 0 the state, bit 1 the diagnostics; the clock always (what = 0:
 * the clock alone, the run loop's clockCopy, src/ui/api.ts:503-534). */
typedef struct shx_readback shx_readback;
enum { SHX_READ_STATE = 1, SHX_READ_DIAG = 2 };
int shx_readback_begin(shx_sim* sim, int what, shx_readback** out);
/* Wait until the readback's copies are done (any thread). */
int shx_readback_wait(shx_readback* rb);
/* After shx_readback_wait: its host data. u1, u2: ncell x 4 values each in the sim's precision
 * (null without SHX_READ_STATE); clock: the 11 clock values; diag: the 24 diagnostics channels
 * (null without SHX_READ_DIAG). Any out-pointer may be null. Valid until shx_readback_free. */
int shx_readback_data(shx_readback* rb, const void** u1, const void** u2, size_t* ncell, int* precision,
                      const double** clock, const double** diag);
int shx_readback_free(shx_readback* rb);

/* ---------------- block AMR (AmrSim, src/amr/amr_sim.ts) ----------------
 * A 2:1-balanced block tree (the reference builds and keeps it in BlockTree,
 * src/amr/tree.ts:106-824; shx_amr_regrid below moves a sim onto a new one): blocks of 16^3 cells with 2
 * ghost layers (tree.ts:14-18), given as (level, ix, iy, iz) per slot, parents and leaves,
 * where block coordinates count blocks of that level from the domain's lower corner.
 * The shx_config describes the level-0 grid: n = 16 * root blocks per axis, dx = level-0 cell
 * size, ndim = 3. State layout: per slot 4096 cells (x fastest) of u1 and u2 as above; only
 * leaves' cells are read and written. ndim = 2 (tree.ts:138-175 with ndim 2): a quadtree of
 * 16^2-cell blocks, root[2] = 1, n[2] = 1, every iz 0, four children per parent (octant
 * x | y << 1), each z face of the table the block itself; 256 cells per slot; the PASSES and
 * LEVELS paths. */
typedef struct shx_amr shx_amr;

/* Kernel path of the AMR step. FUSED (the batched form): a clock pass and one persistent kernel
 * over a device list of leaves that forms ghosts, half steps, faces and the update of each block
 * in one march, coarse-fine faces through a flux register (one launch: workgroups take leaves
 * from a device queue, a leaf beside a refined block after the finer leaves across it). PASSES: the reference's
 * passes (assemble -> clock -> predict -> update, amr_sim.ts:2862-2872) with W, PA and PB atlases
 * in HBM, each launched once over every level from device slot lists. LEVELS: those kernels
 * launched level by level (the plain per-level form). All three give the same results. */
enum { SHX_AMR_PATH_FUSED = 0, SHX_AMR_PATH_PASSES = 1, SHX_AMR_PATH_LEVELS = 2 };

typedef struct {
  int root[3];         /* root blocks per axis (the BlockTree root grid) */
  int path;            /* SHX_AMR_PATH_* */
  int compensated;     /* 1: the four fine fluxes of a coarse face are summed with TwoSum before
                          their mean; 0: plain FP32 sum, as match_faces (amr/update.wgsl:207-290) */
  int time_bins;       /* per-block timesteps (SimConfig.timeBins, amr/bins_on.wgsl): 0 off; K in 1..5:
                          shx_amr_step advances in cycles of 2^K substeps of the fastest block's dt,
                          each leaf stepping 2^k of them at a time (k its bin). PASSES and LEVELS
                          paths only */
  int regrid_every;    /* 0: the tree changes only through shx_amr_regrid; N > 0: every N steps
                          (REGRID_EVERY = 8 in the reference, amr_sim.ts:73, :3020-3060)
                          shx_amr_step measures shx_amr_indicator, forms the next tree with
                          shx_amr_adapt and re-meshes onto it (that step waits for the readback) */
  int max_level;       /* the deepest level such a re-mesh refines to (SimConfig.amrLevels) */
  double refine_above; /* a leaf whose indicator exceeds this splits (SimConfig.refineHi) */
  double coarsen_below;/* a parent whose children are leaves all below this merges (refineLo) */
} shx_amr_layout;

/* AmrSim.create (amr_sim.ts:1924-1992 with uploadTree :2566-2625): validate the tree (levels,
 * coordinates, parents, all-or-no children, 2:1 balance across faces), build the block table
 * and slot lists, allocate state and atlases, capture the step graphs. */
int shx_amr_create(const shx_config* cfg, const shx_amr_layout* layout, const int* blocks, int nblocks,
                   shx_amr** out);
int shx_amr_destroy(shx_amr* sim);
/* BlockTree.table (tree.ts:642-675): nblocks rows of 32 ints: [0] level, [1..3] coordinates,
 * [4] flags (1 active, 2 leaf), [5] parent, [6..11] face neighbours (-x, +x, -y, +y, -z, +z: a
 * same-level slot, -1 where the region is one level coarser, -1 - boundary code at a closed
 * face), [12..19] children, [20..25] the coarser leaf beyond each -1 face; -1 elsewhere. */
int shx_amr_table(shx_amr* sim, int* out, size_t len);
/* AmrSim.setState / readLeaves: ncell = nblocks * 4096 (* 256 in 2D). Resets the clock as shx_set_state. */
int shx_amr_set_state(shx_amr* sim, const float* u1, const float* u2, size_t ncell, double time, int steps);
int shx_amr_set_state_f64(shx_amr* sim, const double* u1, const double* u2, size_t ncell, double time, int steps);
int shx_amr_read_state(shx_amr* sim, float* u1, float* u2);
int shx_amr_read_state_f64(shx_amr* sim, double* u1, double* u2);
/* AmrSim.step / encodeSteps (amr_sim.ts:2709-2947), single global dt (PER_DX clock): enqueue
 * `count` steps; *taken (may be null) = the steps enqueued (`count`). With layout regrid_every
 * = N > 0 every N-th step is followed by a re-mesh (indicator, shx_amr_adapt, shx_amr_regrid),
 * after which the table, lists and slot count are the new tree's (shx_amr_stats, shx_amr_table);
 * otherwise re-meshing is the caller's, through shx_amr_indicator / shx_amr_regrid. */
int shx_amr_step(shx_amr* sim, int count, int* taken);
int shx_amr_sync(shx_amr* sim);
int shx_amr_read_clock(shx_amr* sim, double out[11]);
/* AmrSim.runToEnd (src/amr/amr_sim.ts:3749-3775): as shx_run_to_end on the block grid (the
 * caller re-meshes between runs; batches are sized as shx_run_to_end's). */
int shx_amr_run_to_end(shx_amr* sim, int batch, int max_steps, int* steps);
/* AmrSim.writeParams / buildHydro (src/amr/amr_sim.ts:1677, :1107): as shx_write_params /
 * shx_build_hydro; the tree, boundaries, precision and device of *cfg must equal the sim's. The
 * state, tree and clock are kept. */
int shx_amr_write_params(shx_amr* sim, const shx_config* cfg);
/* Per-block timesteps (layout time_bins > 0): out (may be null; len >= nblocks) the bin of every
 * slot in the last cycle (2^bin substeps a step; a parent its fastest child's); work (may be null)
 * = (cycles, block updates made, deposits included, what single steps would have made: leaves
 * x 2^K per cycle), as AmrSim.binWork (amr_sim.ts:1663-1675). Counted since creation or the
 * last re-mesh. Each cycle reads its rates back once (its bins and lists are planned on the host),
 * so shx_amr_step waits for the previous cycle's steps before enqueueing the next. */
int shx_amr_bins(shx_amr* sim, int* out, size_t len, int64_t work[3]);
int shx_amr_build_hydro(shx_amr* sim, const shx_config* cfg);
/* AmrSim.diagnostics (amr/diagnostics.wgsl:13-117): the channels of shx_diagnostics over leaf
 * cells, each weighted by its own volume h^3 and placed at its centre (blk_pos,
 * amr/grid.wgsl:172-177). */
int shx_amr_diagnostics(shx_amr* sim, double out[24]);
/* A readback of the block grid (as shx_readback_begin): the leaves' state of every slot, the
 * clock and the diagnostics. */
int shx_amr_readback_begin(shx_amr* sim, int what, shx_readback** out);
/* Time `steps` steps pass by pass with HIP events: ms[0] assemble, [1] clock, [2] predict,
 * [3] update, [4] whole step. */
int shx_amr_profile(shx_amr* sim, int steps, double ms[5]);
/* AmrSim.fluxMismatch (src/amr/amr_sim.ts:4078-4160) with DEBUG_FLUX (amr/update.wgsl:324): one
 * step's face mass fluxes as each cell computed them (the state and clock are left as they
 * were). counts[0] same-level faces between blocks, [1] of them differing in any bit, [2]
 * coarse-fine faces; *worst_coarse_fine = the largest |coarse flux - mean of the fine fluxes|. */
int shx_amr_flux_mismatch(shx_amr* sim, int64_t counts[3], double* worst_coarse_fine);
/* Counts: out[0] blocks, [1] leaves, [2] leaves with a refined neighbour, [3] levels,
 * [4] kernel launches per step. */
int shx_amr_stats(shx_amr* sim, int out[5]);
/* The refinement indicator of every leaf (AmrSim's flagging pass, src/amr/amr_sim.ts:3029-3060,
 * with amr/flag.wgsl:66-155 for gas cells): out[slot] = the largest over the leaf's cells of
 * Loehner's normalized second derivative of density and of pressure, neighbours from the
 * assembled ghosts, at least 1e-30 (near 1 at a discontinuity, falling as dx / L on a smooth
 * profile: the scale SimConfig.refineHi / refineLo, 0.8 / 0.2 by default, are set for); NaN for
 * parents. out holds one double per block. */
int shx_amr_indicator(shx_amr* sim, double* out);
/* BlockTree.adapt (src/amr/tree.ts:489-641) on the host, no GPU: the next tree from `blocks`
 * (nblocks x 4 ints, 2:1 balanced, the tree of a sim with root blocks `root` and dimension
 * ndim 2 or 3) and one indicator per block (NaN or any value for parents): leaves above
 * refine_above split (below max_level), parents whose children are leaves all below
 * coarsen_below merge, then 2:1 balance across faces is restored by splitting (periodic[a]
 * != 0: axis a wraps). Writes the blocks, parents first by level then z, y, x, into out (cap
 * rows of 4 ints) and their count into *nout; every block lies within one level of the given
 * tree. SHX_EINVAL if out is too small (then *nout is the count needed). */
int shx_amr_adapt(const int* blocks, int nblocks, const int root[3], int ndim, const double* indicator,
                  double refine_above, double coarsen_below, int max_level, const int periodic[3], int* out,
                  int cap, int* nout);
/* Re-mesh (AmrSim.applyRegrid, src/amr/amr_sim.ts:2648-2697, amr/regrid.wgsl:40-139): replace
 * the tree by `blocks` (nblocks x 4 ints as for shx_amr_create, 2:1 balanced), which must lie
 * within one level of the current tree everywhere: each new leaf is a current leaf (copied), a
 * child of one (prolonged: conservative, limited, the children's mean energy kept) or a current
 * parent of eight leaves, four in 2D (their mean). The clock carries on; table, lists and graphs are rebuilt.
 * On failure the sim is unchanged. */
int shx_amr_regrid(shx_amr* sim, const int* blocks, int nblocks);

/* ---------------- ensembles of independent 2D grids ----------------
 * The reference runs one Sim per device (src/sim.ts); an ensemble advances `count` independent
 * 2D uniform grids, each its own size and clock (dt per instance), in one persistent launch per
 * step over a device list of column tiles from every instance. Each instance is the run a Sim
 * with the same configuration would make (the plain path's arithmetic); cfg gives what they
 * share (limiter, solver, gamma, cfl, boundaries, floors, precision); sizes[2i], sizes[2i+1] are
 * instance i's nx, ny on the unit box (dx = 1 / nx). Boundaries: each of x and y periodic on both
 * faces, or outflow / reflecting (outflow-only faces and an axis periodic on one face only are
 * refused with SHX_EINVAL: a Sim runs those on its plain path). */
typedef struct shx_ens shx_ens;
int shx_ens_create(const shx_config* cfg, int count, const int* sizes, shx_ens** out);
int shx_ens_destroy(shx_ens* ens);
/* Instance i's state, nx*ny cells of u1 and u2 in the Sim layout; resets its clock. */
int shx_ens_set_state(shx_ens* ens, int i, const float* u1, const float* u2, size_t ncell);
int shx_ens_set_state_f64(shx_ens* ens, int i, const double* u1, const double* u2, size_t ncell);
int shx_ens_read_state(shx_ens* ens, int i, float* u1, float* u2);
int shx_ens_read_state_f64(shx_ens* ens, int i, double* u1, double* u2);
/* `count` steps of every instance (graph launches). */
int shx_ens_step(shx_ens* ens, int count);
int shx_ens_sync(shx_ens* ens);
int shx_ens_read_clock(shx_ens* ens, int i, double out[11]);
/* ms[0] clock pass, ms[1] the batched kernel, ms[2] whole step (HIP events, averaged). */
int shx_ens_profile(shx_ens* ens, int steps, double ms[3]);
/* Time the tile candidates on the current states (which it restores); chosen = (tx, ty). */
int shx_ens_autotune(shx_ens* ens, int reps, int chosen[2]);

/* The calling thread's last error message ("" if none). */
const char* shx_last_error(void);
/* Library version and the device it sees: "shx <ver> gfx950 <name>" */
const char* shx_version(void);

#ifdef __cplusplus
}
#endif
#endif
