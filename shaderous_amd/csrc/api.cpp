// C-ABI of libshx for the uniform grid (include/shx.h): the sim object, its device memory,
// stream and captured step graphs. Every entry catches its own failures and returns a status;
// nothing throws across the ABI.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/shx.h"
#include "common.hpp"
#include "ens_impl.hpp"
#include "uniform.hpp"

using namespace shx;

namespace shx {
thread_local std::string g_err;
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace shx

struct shx_sim {
  shx_config cfg;
  int es;                    // element size: 4 or 8
  size_t ncell;
  hipStream_t stream = nullptr;
  void* U1[2] = {nullptr, nullptr};
  void* U2[2] = {nullptr, nullptr};
  void* PA = nullptr;
  void* PB = nullptr;
  // the clock block (clock_block_bytes): the clocks of parity 0 and 1 (16 T each: the Clock's 11
  // values and spare), three rate slots, two epochs. The plain path keeps its clock in the
  // parity-0 clock and the rate in slot 0; the fused step rotates them (StepClock).
  void* clk = nullptr;
  unsigned long long* slot = nullptr;   // in the clock block
  unsigned int* epoch = nullptr;        // in the clock block
  double* diag_part = nullptr;           // diag_blocks records of DIAG_CH channels
  double* diag_out = nullptr;            // DIAG_CH channels
  int diag_blocks = 256;
  int cur = 0;
  bool fused = false;
  bool zbc = false;          // fused path with a non-periodic z face: PA, PB hold planes 0 and nz - 1
  int tile = FUSED_32x16;
  int zc = 32;
  Params<float> pf;
  Params<double> pd;
  UniformOps ops{};
  // [parity][1, 2 or 8 steps]: a step of a small grid is shorter than a graph launch's host
  // cost, so runs of steps go out eight to a launch
  hipGraphExec_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  // A flat grid on the fused path steps through the 2D kernels as a batch of one grid: in FP32
  // the row march (k_ens_rows, one wave per 60-cell strip) or the LDS tile kernel (k_ens_fused),
  // by tile; in FP64 the tile kernel's 16 x 16 form. The 3D march (ops.fused) remains its
  // flux-debug form.
  bool flat2d = false;
  EnsOps eops{};
  void* einst = nullptr;                  // one EnsInst<T>: this grid's parameters, offset 0
  void* etiles[ENS_NTILES] = {};          // the grid's tiles (strips) for every 2D tile shape
  int entiles[ENS_NTILES] = {};
  int etile = ENS_ROWS16;
  int egrid = 1024;                       // persistent grid of the tile kernel
};

namespace {

Launch launch_of(shx_sim* s) {
  Launch L;
  L.stream = s->stream;
  L.params = s->es == 8 ? (const void*)&s->pd : (const void*)&s->pf;
  L.ncell = s->ncell;
  return L;
}

// The clocks of the fused step from parity p (its clock pass is in the step kernel).
ClockRefs clock_refs(shx_sim* s, int p) { return ClockRefs{s->clk, s->slot, s->epoch, p}; }

// The clock the next step reads: the fused path's of the current parity, the plain path's one.
const void* current_clock(const shx_sim* s) {
  return static_cast<const char*>(s->clk) + (s->fused ? (size_t)16 * s->es * s->cur : 0);
}

// The fused path's one-pass step from parity p into 1 - p (its clock pass included; after the
// boundary planes' half steps): the 2D kernels on a flat FP32 grid, else the 3D march with tile
// `tile`.
void enqueue_fused(shx_sim* s, int p, int tile, int zc, bool debug) {
  if (s->flat2d && !debug) {
    s->eops.fused(s->stream, tile, s->egrid, s->einst, 1, s->etiles[tile], s->entiles[tile], s->U1[p], s->U2[p],
                  s->U1[1 - p], s->U2[1 - p], clock_refs(s, p), s->es == 8 ? (const void*)&s->pd : (const void*)&s->pf);
    return;
  }
  s->ops.fused(launch_of(s), tile, s->U1[p], s->U2[p], s->U1[1 - p], s->U2[1 - p], clock_refs(s, p),
               s->zbc ? s->PA : nullptr, s->zbc ? s->PB : nullptr, zc, debug);
}

// One step from parity p into 1 - p.
void enqueue_step(shx_sim* s, int p, bool debug = false) {
  Launch L = launch_of(s);
  const UniformOps& o = s->ops;
  if (s->fused) {
    if (s->zbc) o.zplanes(L, s->U1[p], s->U2[p], s->PA, s->PB, clock_refs(s, p));
    enqueue_fused(s, p, s->flat2d && !debug ? s->etile : s->tile, s->zc, debug);
  } else {
    o.rates(L, s->U1[p], s->U2[p], s->slot);
    o.clock(L, s->clk, s->slot);
    o.predict(L, s->U1[p], s->U2[p], s->PA, s->PB, s->clk);
    o.update(L, s->U1[p], s->U2[p], s->PA, s->PB, s->U1[1 - p], s->U2[1 - p], s->clk, debug);
  }
}

void destroy_graphs(shx_sim* s) {
  for (int p = 0; p < 2; p++)
    for (int k = 0; k < 3; k++)
      if (s->graph[p][k]) { (void)hipGraphExecDestroy(s->graph[p][k]); s->graph[p][k] = nullptr; }
}

int build_graphs(shx_sim* s) {
  destroy_graphs(s);
  for (int p = 0; p < 2; p++)
    for (int k = 0; k < 3; k++) {
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      const int n = k == 0 ? 1 : (k == 1 ? 2 : 8);
      for (int i = 0; i < n; i++) enqueue_step(s, (p + i) & 1);
      HIPCHK(hipStreamEndCapture(s->stream, &g));
      HIPCHK(hipGraphInstantiate(&s->graph[p][k], g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
    }
  return SHX_OK;
}

// The rate slots: the plain path recomputes the max rate each step (k_rates into slot 0); the
// fused path's step leaves the rate of the state it writes in the next epoch's slot; after a
// state change the epochs restart at 0 and the current state's rate is recomputed into slot 0.
int reset_rates(shx_sim* s) {
  HIPCHK(hipMemsetAsync(s->slot, 0, 3 * sizeof(unsigned long long) + 2 * sizeof(unsigned int), s->stream));
  if (s->fused) {
    Launch L = launch_of(s);
    s->ops.rates(L, s->U1[s->cur], s->U2[s->cur], s->slot);
  }
  return SHX_OK;
}

bool fused_supported(const shx_config& c) {
  // z: any face of a 3D grid (a non-periodic one through the boundary planes' half steps), a
  // flat grid periodic; x/y: periodic, outflow, reflecting
  if (c.n[2] < 2 && (c.bc_lo[2] != SHX_BC_PERIODIC || c.bc_hi[2] != SHX_BC_PERIODIC)) return false;
  // an x or y axis periodic on one face only: its ghost columns' half steps would need the
  // wrapped cells' own stencils (the plain path takes them)
  for (int a = 0; a < 2; a++)
    if ((c.bc_lo[a] == SHX_BC_PERIODIC) != (c.bc_hi[a] == SHX_BC_PERIODIC)) return false;
  for (int a = 0; a < 2; a++)
    if (c.bc_lo[a] == SHX_BC_OUTFLOW_ONLY || c.bc_hi[a] == SHX_BC_OUTFLOW_ONLY) return false;
  for (int a = 0; a < 2; a++)
    if (c.n[a] < 2) return false;
  return true;
}

template <typename T>
int set_state_t(shx_sim* s, const T* u1, const T* u2, size_t ncell, double time, int steps) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if ((int)sizeof(T) != s->es) return fail(SHX_EINVAL, "state precision %d does not match the sim's %d", (int)sizeof(T) * 8, s->es * 8);
  if (ncell != s->ncell) return fail(SHX_EINVAL, "ncell %zu != %zu", ncell, s->ncell);
  size_t bytes = ncell * 4 * sizeof(T);
  for (size_t o = 0; o < 4 * ncell; o++)
    if (!isfinite((double)u1[o]) || !isfinite((double)u2[o])) return fail(SHX_ESTATE, "non-finite state at %zu", o / 4);
  s->cur = 0;
  HIPCHK(hipMemcpyAsync(s->U1[0], u1, bytes, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->U2[0], u2, bytes, hipMemcpyHostToDevice, s->stream));
  T clk[16] = {0};
  clk[1] = T(time);
  clk[3] = T(steps);
  clk[4] = clk[5] = clk[6] = T(1);
  HIPCHK(hipMemcpyAsync(s->clk, clk, sizeof clk, hipMemcpyHostToDevice, s->stream));
  int rc = reset_rates(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return SHX_OK;
}

template <typename T>
int read_state_t(shx_sim* s, T* u1, T* u2) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if ((int)sizeof(T) != s->es) return fail(SHX_EINVAL, "state precision mismatch");
  size_t bytes = s->ncell * 4 * sizeof(T);
  HIPCHK(hipMemcpyAsync(u1, s->U1[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(u2, s->U2[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return SHX_OK;
}

// The kernels of a configuration's scheme (limiter, Riemann solver, reconstruction): the 3D
// tables, and for a flat FP32 grid whose tile lists exist the 2D kernels. False if not built.
bool select_kernels(shx_sim* s, const shx_config& c, bool lists) {
  UniformOps ops = uniform_ops(c.precision, c.limiter, c.riemann, c.isentropic ? 1 : 0);
  if (!ops.rates) return false;
  s->ops = ops;
  s->flat2d = false;
  if (lists && c.n[2] == 1) {
    const int isen = c.isentropic ? 1 : 0;
    s->eops = s->es == 8 ? ens_ops_f64(c.limiter, c.riemann, isen) : ens_ops_f32(c.limiter, c.riemann, isen);
    s->flat2d = s->eops.fused != nullptr;
  }
  return true;
}

// New parameter blocks from a configuration (run-time fields), on the device where a kernel
// reads them from memory (the 2D kernels' instance record).
int apply_params(shx_sim* s, const shx_config& c) {
  s->cfg = c;
  s->pf = make_params<float>(c);
  s->pd = make_params<double>(c);
  if (s->einst) {
    const EnsInst<float> one{s->pf, 0};
    const EnsInst<double> one64{s->pd, 0};
    if (s->es == 8) HIPCHK(hipMemcpyAsync(s->einst, &one64, sizeof one64, hipMemcpyHostToDevice, s->stream));
    else HIPCHK(hipMemcpyAsync(s->einst, &one, sizeof one, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return SHX_OK;
}

// What a running sim cannot change: the grid and its storage, the boundaries, the precision,
// the path and the device (and, unless `scheme`, the limiter, solver and reconstruction).
bool same_layout(const shx_config& a, const shx_config& b, bool scheme) {
  for (int k = 0; k < 3; k++)
    if (a.n[k] != b.n[k] || a.bc_lo[k] != b.bc_lo[k] || a.bc_hi[k] != b.bc_hi[k]) return false;
  if (a.ndim != b.ndim || a.precision != b.precision || a.path != b.path || a.device != b.device || a.dx != b.dx)
    return false;
  if (!scheme && (a.limiter != b.limiter || a.riemann != b.riemann || (a.isentropic != 0) != (b.isentropic != 0)))
    return false;
  return true;
}

int reconfigure(shx_sim* s, const shx_config* cfg, bool scheme) {
  if (!s || !cfg) return fail(SHX_EINVAL, "null argument");
  if (!same_layout(s->cfg, *cfg, scheme))
    return fail(SHX_EINVAL, scheme ? "the grid, boundaries, precision, path and device of a sim are fixed"
                                   : "the grid, boundaries, precision, path, device and scheme of a sim are fixed "
                                     "(shx_build_hydro changes the scheme)");
  if (!(cfg->gamma > 1) || !(cfg->cfl > 0)) return fail(SHX_EINVAL, "gamma must exceed 1 and cfl be positive");
  shx_config c = *cfg;
  for (int k = 0; k < 3; k++) c.tile[k] = s->cfg.tile[k];
  HIPCHK(hipStreamSynchronize(s->stream));
  if (scheme) {
    shx_sim probe = {};
    probe.es = s->es;
    if (!select_kernels(&probe, c, false))
      return fail(SHX_EINVAL, "no kernels built for precision %d limiter %d riemann %d isentropic %d", c.precision,
                  c.limiter, c.riemann, c.isentropic);
    select_kernels(s, c, s->etiles[0] != nullptr);
  }
  int rc = apply_params(s, c);
  if (rc) return rc;
  // the rate of the current state depends on gamma and ndim
  rc = reset_rates(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return build_graphs(s);
}

}  // namespace

namespace shx {
int readback_begin(hipStream_t stream, int what, int es, size_t ncell, const void* U1, const void* U2, const void* clk,
                   const double* diag, shx_readback** out) {
  *out = nullptr;
  shx_readback* r = new shx_readback();
  r->ncell = (what & SHX_READ_STATE) ? ncell : 0;
  r->es = es;
  r->what = what;
  const size_t sb = r->ncell * 4 * es;
  const size_t bytes = 2 * sb + 16 * 8 + DIAG_CH * 8;
  if (hipHostMalloc((void**)&r->host, bytes, hipHostMallocDefault) != hipSuccess) {
    delete r;
    return fail(SHX_ENOMEM, "pinned host allocation of %zu bytes for a readback failed", bytes);
  }
  hipError_t e = hipEventCreateWithFlags(&r->ev, hipEventDisableTiming);
  if (e == hipSuccess && sb) e = hipMemcpyAsync(r->host, U1, sb, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && sb) e = hipMemcpyAsync(r->host + sb, U2, sb, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(r->host + 2 * sb, clk, 16 * es, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && (what & SHX_READ_DIAG))
    e = hipMemcpyAsync(r->host + 2 * sb + 16 * 8, diag, DIAG_CH * 8, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipEventRecord(r->ev, stream);
  if (e != hipSuccess) {
    // nothing may still write into the host memory once it is freed
    (void)hipStreamSynchronize(stream);
    shx_readback_free(r);
    return fail(SHX_EDEVICE, "enqueueing a readback: %s", hipGetErrorString(e));
  }
  *out = r;
  return SHX_OK;
}
}  // namespace shx

extern "C" {

int shx_readback_wait(shx_readback* r) {
  if (!r) return fail(SHX_EINVAL, "null readback");
  if (r->done) return SHX_OK;
  HIPCHK(hipEventSynchronize(r->ev));
  const size_t sb = r->ncell * 4 * r->es;
  const unsigned char* c = r->host + 2 * sb;
  for (int i = 0; i < 11; i++) r->clock[i] = r->es == 8 ? ((const double*)c)[i] : (double)((const float*)c)[i];
  if (r->what & SHX_READ_DIAG) memcpy(r->diag, r->host + 2 * sb + 16 * 8, sizeof r->diag);
  r->done = true;
  return SHX_OK;
}

int shx_readback_data(shx_readback* r, const void** u1, const void** u2, size_t* ncell, int* precision,
                      const double** clock, const double** diag) {
  if (!r) return fail(SHX_EINVAL, "null readback");
  if (!r->done) return fail(SHX_EINVAL, "the readback has not been waited for");
  const size_t sb = r->ncell * 4 * r->es;
  if (u1) *u1 = r->ncell ? r->host : nullptr;
  if (u2) *u2 = r->ncell ? r->host + sb : nullptr;
  if (ncell) *ncell = r->ncell;
  if (precision) *precision = r->es * 8;
  if (clock) *clock = r->clock;
  if (diag) *diag = (r->what & SHX_READ_DIAG) ? r->diag : nullptr;
  return SHX_OK;
}

int shx_readback_free(shx_readback* r) {
  if (!r) return SHX_OK;
  // a readback freed before it was waited for: its copies still target the host memory
  if (!r->done && r->ev) (void)hipEventSynchronize(r->ev);
  if (r->ev) (void)hipEventDestroy(r->ev);
  if (r->host) (void)hipHostFree(r->host);
  delete r;
  return SHX_OK;
}

int shx_readback_begin(shx_sim* s, int what, shx_readback** out) {
  if (!s || !out) return fail(SHX_EINVAL, "null argument");
  if (what & SHX_READ_DIAG) {
    Launch L = launch_of(s);
    s->ops.diag(L, uniform_diag_geo(s->cfg), s->U1[s->cur], s->U2[s->cur], s->diag_part, s->diag_blocks, s->diag_out);
    HIPCHK(hipGetLastError());
  }
  return readback_begin(s->stream, what, s->es, s->ncell, s->U1[s->cur], s->U2[s->cur], current_clock(s), s->diag_out,
                        out);
}

int shx_write_params(shx_sim* s, const shx_config* cfg) { return reconfigure(s, cfg, false); }
int shx_build_hydro(shx_sim* s, const shx_config* cfg) { return reconfigure(s, cfg, true); }

const char* shx_last_error(void) { return g_err.c_str(); }

const char* shx_version(void) {
  static std::string v;
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    v = std::string("shx 0.1 ") + prop.gcnArchName + " " + prop.name;
  else
    v = "shx 0.1 (no device)";
  return v.c_str();
}

void shx_config_default(shx_config* c, int nx, int ny, int nz) {
  memset(c, 0, sizeof *c);
  c->n[0] = nx; c->n[1] = ny; c->n[2] = nz;
  c->ndim = nz == 1 ? 2 : 3;
  c->limiter = SHX_LIM_KOREN;
  c->riemann = SHX_RS_HLLC;
  c->isentropic = 1;
  c->precision = 32;
  c->path = SHX_PATH_FUSED;
  c->dx = 1.0 / nx;
  c->gamma = 1.4;
  c->cfl = 0.8;
  c->rho_floor = 1e-10;
  c->p_floor = 1e-14;
  c->eta_dual = 1e-3;
  c->shock_thresh = 0.5;
  c->t_end = 1e30;
  for (int a = 0; a < 3; a++) c->org[a] = -0.5 * nx * c->dx;
}

int shx_create(const shx_config* cfg, shx_sim** out) {
  if (!cfg || !out) return fail(SHX_EINVAL, "null argument");
  *out = nullptr;
  const shx_config& c = *cfg;
  for (int a = 0; a < 3; a++) {
    if (c.n[a] < 1) return fail(SHX_EINVAL, "n[%d] = %d", a, c.n[a]);
    if (c.bc_lo[a] < 0 || c.bc_lo[a] > 3 || c.bc_hi[a] < 0 || c.bc_hi[a] > 3) return fail(SHX_EINVAL, "boundary code on axis %d", a);
  }
  if (c.n[2] == 1 && (c.bc_lo[2] != 0 || c.bc_hi[2] != 0)) return fail(SHX_EINVAL, "a single layer along z must be periodic");
  if (c.precision != 32 && c.precision != 64) return fail(SHX_EINVAL, "precision %d", c.precision);
  if (!(c.gamma > 1) || !(c.dx > 0) || !(c.cfl > 0)) return fail(SHX_EINVAL, "gamma, dx and cfl must be positive (gamma > 1)");
  UniformOps ops = uniform_ops(c.precision, c.limiter, c.riemann, c.isentropic ? 1 : 0);
  if (!ops.rates) return fail(SHX_EINVAL, "no kernels built for precision %d limiter %d riemann %d isentropic %d", c.precision, c.limiter, c.riemann, c.isentropic);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c.device) return fail(SHX_EDEVICE, "no HIP device %d", c.device);
  HIPCHK(hipSetDevice(c.device));

  shx_sim* s = new shx_sim();
  s->cfg = c;
  s->es = c.precision == 64 ? 8 : 4;
  s->ncell = (size_t)c.n[0] * c.n[1] * c.n[2];
  s->ops = ops;
  s->pf = make_params<float>(c);
  s->pd = make_params<double>(c);
  s->fused = c.path == SHX_PATH_FUSED && fused_supported(c);
  s->tile = s->es == 8 ? FUSED_20x20 : FUSED_32x16;
  s->zc = 32;
  if (c.tile[0] > 0) {
    for (int k = 0; k < FUSED_NTILES; k++) {
      int tx, ty;
      fused_tile_cells(k, tx, ty);
      if (tx == c.tile[0] && ty == c.tile[1]) s->tile = k;
    }
    if (s->es == 8) s->tile = FUSED_20x20;
  }
  if (c.tile[2] > 0) s->zc = c.tile[2];
  if (s->zc > c.n[2]) s->zc = c.n[2];
  if (s->fused && c.n[2] == 1) {
    const int isen = c.isentropic ? 1 : 0;
    s->eops = s->es == 8 ? ens_ops_f64(c.limiter, c.riemann, isen) : ens_ops_f32(c.limiter, c.riemann, isen);
    s->flat2d = s->eops.fused != nullptr;
    if (s->es == 8) s->etile = ENS_20x20;   // the FP64 2D kernel's one tile
  }
  if (s->flat2d && s->es == 4 && c.tile[0] > 0)
    for (int k = 0; k < ENS_NTILES; k++) {
      int tx, ty;
      ens_tile_cells(k, tx, ty);
      if (tx == c.tile[0] && ty == c.tile[1]) s->etile = k;
    }

  size_t bytes = s->ncell * 4 * s->es;
  auto alloc = [&](void** p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
  bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
  for (int p = 0; ok && p < 2; p++) ok = alloc(&s->U1[p], bytes) && alloc(&s->U2[p], bytes);
  s->zbc = s->fused && (c.bc_lo[2] != SHX_BC_PERIODIC || c.bc_hi[2] != SHX_BC_PERIODIC);
  if (ok && !s->fused) ok = alloc(&s->PA, bytes) && alloc(&s->PB, bytes);
  if (ok && s->zbc) {
    const size_t zb = (size_t)2 * c.n[0] * c.n[1] * 4 * s->es;
    ok = alloc(&s->PA, zb) && alloc(&s->PB, zb);
  }
  ok = ok && alloc(&s->clk, clock_block_bytes(1, s->es)) && alloc((void**)&s->diag_part, (size_t)s->diag_blocks * DIAG_CH * 8) &&
       alloc((void**)&s->diag_out, DIAG_CH * 8);
  std::vector<EnsTile> lists[ENS_NTILES];
  if (s->flat2d) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus < 1) cus = 256;
    s->egrid = 4 * cus;
    ok = ok && alloc(&s->einst, s->es == 8 ? sizeof(EnsInst<double>) : sizeof(EnsInst<float>));
    for (int k = 0; k < ENS_NTILES; k++) {
      int tx, ty;
      ens_tile_cells(k, tx, ty);
      for (int y0 = 0; y0 < c.n[1]; y0 += ty)
        for (int x0 = 0; x0 < c.n[0]; x0 += tx) lists[k].push_back(EnsTile{0, x0, y0, std::min(y0 + ty, c.n[1])});
      s->entiles[k] = (int)lists[k].size();
      ok = ok && alloc(&s->etiles[k], lists[k].size() * sizeof(EnsTile));
    }
  }
  if (!ok) {
    shx_destroy(s);
    return fail(SHX_ENOMEM, "device allocation failed for %zu cells (%.2f GB per state copy)", s->ncell, bytes * 2 / 1e9);
  }
  (void)hipMemset(s->clk, 0, clock_block_bytes(1, s->es));
  s->slot = reinterpret_cast<unsigned long long*>(static_cast<char*>(s->clk) + clock_slots_at(1, s->es));
  s->epoch = reinterpret_cast<unsigned int*>(static_cast<char*>(s->clk) + clock_epochs_at(1, s->es));
  if (s->flat2d) {
    const EnsInst<float> one{s->pf, 0};
    const EnsInst<double> one64{s->pd, 0};
    bool copied = s->es == 8 ? hipMemcpy(s->einst, &one64, sizeof one64, hipMemcpyHostToDevice) == hipSuccess
                             : hipMemcpy(s->einst, &one, sizeof one, hipMemcpyHostToDevice) == hipSuccess;
    for (int k = 0; k < ENS_NTILES; k++)
      copied = copied && hipMemcpy(s->etiles[k], lists[k].data(), lists[k].size() * sizeof(EnsTile), hipMemcpyHostToDevice) == hipSuccess;
    if (!copied) {
      shx_destroy(s);
      return fail(SHX_EDEVICE, "copying the grid's tile lists to the device failed");
    }
  }
  int rc = build_graphs(s);
  if (rc) { shx_destroy(s); return rc; }
  *out = s;
  return SHX_OK;
}

int shx_destroy(shx_sim* s) {
  if (!s) return SHX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  destroy_graphs(s);
  for (int p = 0; p < 2; p++) { (void)hipFree(s->U1[p]); (void)hipFree(s->U2[p]); }
  (void)hipFree(s->PA); (void)hipFree(s->PB); (void)hipFree(s->clk);
  (void)hipFree(s->diag_part); (void)hipFree(s->diag_out);
  (void)hipFree(s->einst);
  for (int k = 0; k < ENS_NTILES; k++) (void)hipFree(s->etiles[k]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return SHX_OK;
}

int shx_set_state(shx_sim* s, const float* u1, const float* u2, size_t ncell, double time, int steps) {
  return set_state_t<float>(s, u1, u2, ncell, time, steps);
}
int shx_set_state_f64(shx_sim* s, const double* u1, const double* u2, size_t ncell, double time, int steps) {
  return set_state_t<double>(s, u1, u2, ncell, time, steps);
}
int shx_read_state(shx_sim* s, float* u1, float* u2) { return read_state_t<float>(s, u1, u2); }
int shx_read_state_f64(shx_sim* s, double* u1, double* u2) { return read_state_t<double>(s, u1, u2); }

int shx_step(shx_sim* s, int count, int* taken) {
  if (taken) *taken = 0;
  if (!s || count < 0) return fail(SHX_EINVAL, "bad argument");
  const int want = count;
  while (count >= 8) {
    HIPCHK(hipGraphLaunch(s->graph[s->cur][2], s->stream));
    count -= 8;
  }
  while (count >= 2) {
    HIPCHK(hipGraphLaunch(s->graph[s->cur][1], s->stream));
    count -= 2;
  }
  if (count == 1) {
    HIPCHK(hipGraphLaunch(s->graph[s->cur][0], s->stream));
    s->cur = 1 - s->cur;
  }
  if (taken) *taken = want;
  return SHX_OK;
}

int shx_sync(shx_sim* s) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  HIPCHK(hipStreamSynchronize(s->stream));
  return SHX_OK;
}

int shx_read_clock(shx_sim* s, double out[11]) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  unsigned char buf[16 * 8];
  HIPCHK(hipMemcpyAsync(buf, current_clock(s), 16 * s->es, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int i = 0; i < 11; i++) out[i] = s->es == 8 ? ((double*)buf)[i] : (double)((float*)buf)[i];
  return SHX_OK;
}

int shx_run_to_end(shx_sim* s, int batch, int max_steps, int* steps) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if (!(s->cfg.t_end < 1e30)) return fail(SHX_EINVAL, "run_to_end needs a finite t_end");
  return run_to_end_loop(
      [s](int n) { return shx_step(s, n, nullptr); }, [s](double* c) { return shx_read_clock(s, c); }, s->cfg.t_end,
      batch, max_steps, steps);
}

int shx_diagnostics(shx_sim* s, double out[24]) {
  if (!s || !out) return fail(SHX_EINVAL, "bad argument");
  Launch L = launch_of(s);
  s->ops.diag(L, uniform_diag_geo(s->cfg), s->U1[s->cur], s->U2[s->cur], s->diag_part, s->diag_blocks, s->diag_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->diag_out, DIAG_CH * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return SHX_OK;
}

int shx_flux_mismatch(shx_sim* s, int64_t* mismatched, int64_t* faces) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  // one step in DEBUG_FLUX mode: the face masses land in the other parity's state arrays; the
  // clock advances as for a step, so a copy of it is restored afterwards
  std::vector<unsigned char> saved(clock_block_bytes(1, s->es));
  HIPCHK(hipMemcpyAsync(saved.data(), s->clk, saved.size(), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  int p = s->cur;
  enqueue_step(s, p, true);
  HIPCHK(hipGetLastError());
  size_t nc = s->ncell;
  std::vector<double> a(4 * nc), b(4 * nc);
  if (s->es == 8) {
    HIPCHK(hipMemcpyAsync(a.data(), s->U1[1 - p], 32 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(b.data(), s->U2[1 - p], 32 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    std::vector<float> fa(4 * nc), fb(4 * nc);
    HIPCHK(hipMemcpyAsync(fa.data(), s->U1[1 - p], 16 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(fb.data(), s->U2[1 - p], 16 * nc, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < 4 * nc; i++) { a[i] = fa[i]; b[i] = fb[i]; }
  }
  HIPCHK(hipMemcpyAsync(s->clk, saved.data(), saved.size(), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int* n = s->cfg.n;
  int64_t mis = 0, tot = 0;
  auto lo = [&](size_t o, int axis) { return axis < 2 ? a[4 * o + 2 * axis] : b[4 * o]; };
  auto hi = [&](size_t o, int axis) { return axis < 2 ? a[4 * o + 2 * axis + 1] : b[4 * o + 1]; };
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        size_t o = ((size_t)k * n[1] + j) * n[0] + i;
        for (int axis = 0; axis < (n[2] == 1 ? 2 : 3); axis++) {
          int c[3] = {i, j, k};
          c[axis] += 1;
          if (c[axis] >= n[axis]) {
            if (s->cfg.bc_hi[axis] != SHX_BC_PERIODIC) continue;
            c[axis] = 0;
          }
          size_t q = ((size_t)c[2] * n[1] + c[1]) * n[0] + c[0];
          double x = hi(o, axis), y = lo(q, axis);
          tot++;
          if (memcmp(&x, &y, sizeof x) != 0) mis++;
        }
      }
  if (mismatched) *mismatched = mis;
  if (faces) *faces = tot;
  return SHX_OK;
}

int shx_profile(shx_sim* s, int steps, double ms[5]) {
  if (!s || !ms || steps < 1) return fail(SHX_EINVAL, "bad argument");
  hipEvent_t ev[6];
  for (int i = 0; i < 6; i++) HIPCHK(hipEventCreate(&ev[i]));
  double acc[5] = {0, 0, 0, 0, 0};
  Launch L = launch_of(s);
  for (int n = 0; n < steps; n++) {
    int p = s->cur;
    HIPCHK(hipEventRecord(ev[0], s->stream));
    if (s->fused) {
      // the clock pass is inside the step kernel: ms[1] stays 0
      HIPCHK(hipEventRecord(ev[1], s->stream));
      HIPCHK(hipEventRecord(ev[2], s->stream));
      if (s->zbc) s->ops.zplanes(L, s->U1[p], s->U2[p], s->PA, s->PB, clock_refs(s, p));
      HIPCHK(hipEventRecord(ev[3], s->stream));
      enqueue_fused(s, p, s->flat2d ? s->etile : s->tile, s->zc, false);
    } else {
      s->ops.rates(L, s->U1[p], s->U2[p], s->slot);
      HIPCHK(hipEventRecord(ev[1], s->stream));
      s->ops.clock(L, s->clk, s->slot);
      HIPCHK(hipEventRecord(ev[2], s->stream));
      s->ops.predict(L, s->U1[p], s->U2[p], s->PA, s->PB, s->clk);
      HIPCHK(hipEventRecord(ev[3], s->stream));
      s->ops.update(L, s->U1[p], s->U2[p], s->PA, s->PB, s->U1[1 - p], s->U2[1 - p], s->clk, false);
    }
    HIPCHK(hipEventRecord(ev[4], s->stream));
    HIPCHK(hipEventSynchronize(ev[4]));
    s->cur = 1 - p;
    float t;
    for (int i = 0; i < 4; i++) { HIPCHK(hipEventElapsedTime(&t, ev[i], ev[i + 1])); acc[i] += t; }
    HIPCHK(hipEventElapsedTime(&t, ev[0], ev[4]));
    acc[4] += t;
  }
  for (int i = 0; i < 5; i++) ms[i] = acc[i] / steps;
  for (int i = 0; i < 6; i++) (void)hipEventDestroy(ev[i]);
  return SHX_OK;
}

int shx_kernel_name(shx_sim* s, char* buf, size_t len) {
  if (!s || !buf || len == 0) return fail(SHX_EINVAL, "bad argument");
  // threads by role: the FP32 tiles 32 x 16, 32 x 8 and 16 x 16
  const bool roles = s->es == 4 && (s->tile == FUSED_R32x16 || s->tile == FUSED_R32x8 || s->tile == FUSED_20x20);
  const bool rows = s->etile == ENS_ROWS8 || s->etile == ENS_ROWS16 || s->etile == ENS_ROWS32;
  const char* fused = s->flat2d ? (rows ? "k_ens_rows" : "k_ens_fused") : (roles ? "k_fused_r" : "k_fused");
  snprintf(buf, len, "%s", !s->fused ? "k_update" : fused);
  return SHX_OK;
}

int shx_autotune(shx_sim* s, int reps, int chosen[3]) {
  if (!s) return fail(SHX_EINVAL, "null sim");
  if (!s->fused || s->es == 8) {
    if (chosen) { chosen[0] = 0; chosen[1] = 0; chosen[2] = s->zc; }
    return SHX_OK;
  }
  if (reps < 1) reps = 3;
  // keep the state: each candidate takes `reps` steps from the current state (its clock and rate
  // slots too), which is restored before the next and at the end
  size_t bytes = s->ncell * 16;
  void *b1 = nullptr, *b2 = nullptr;
  HIPCHK(hipMalloc(&b1, bytes));
  HIPCHK(hipMalloc(&b2, bytes));
  HIPCHK(hipMemcpyAsync(b1, s->U1[s->cur], bytes, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(b2, s->U2[s->cur], bytes, hipMemcpyDeviceToDevice, s->stream));
  std::vector<unsigned char> clk_saved(clock_block_bytes(1, s->es));
  HIPCHK(hipMemcpyAsync(clk_saved.data(), s->clk, clk_saved.size(), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int start = s->cur;
  auto restore = [&]() -> int {
    HIPCHK(hipMemcpyAsync(s->U1[start], b1, bytes, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->U2[start], b2, bytes, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->clk, clk_saved.data(), clk_saved.size(), hipMemcpyHostToDevice, s->stream));
    return SHX_OK;
  };
  const int tiles3[] = {FUSED_R32x16, FUSED_R32x8, FUSED_32x32, FUSED_32x16, FUSED_20x20};
  const int tiles2[] = {ENS_32x32, ENS_32x16, ENS_20x20, ENS_ROWS8, ENS_ROWS16, ENS_ROWS32};
  std::vector<int> tiles = s->flat2d ? std::vector<int>(tiles2, tiles2 + 6) : std::vector<int>(tiles3, tiles3 + 5);
  const int zcs[] = {16, 32, 64, 128};
  double best = 1e30;
  int bt = s->flat2d ? s->etile : s->tile, bz = s->zc;
  Launch L = launch_of(s);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int ti : tiles)
    for (int z : zcs) {
      // the 2D kernels have no plane chunk: one candidate per tile; a chunk beyond the grid's
      // planes is the whole grid, timed once
      if (s->flat2d ? z != zcs[0] : (z > s->cfg.n[2] && z != zcs[0])) continue;
      int rc = restore();
      if (rc) return rc;
      // a warm step, then `reps` timed steps (each with its boundary planes' half steps)
      for (int r = 0, p = start; r <= reps; r++, p = 1 - p) {
        if (r == 1) HIPCHK(hipEventRecord(e0, s->stream));
        if (s->zbc) s->ops.zplanes(L, s->U1[p], s->U2[p], s->PA, s->PB, clock_refs(s, p));
        enqueue_fused(s, p, ti, z, false);
      }
      HIPCHK(hipEventRecord(e1, s->stream));
      HIPCHK(hipEventSynchronize(e1));
      HIPCHK(hipGetLastError());
      float t;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      if (t < best) { best = t; bt = ti; bz = z; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s->flat2d) {
    s->etile = bt;
  } else {
    s->tile = bt;
    s->zc = std::min(bz, s->cfg.n[2]);
  }
  int rc = restore();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->cur = start;
  (void)hipFree(b1);
  (void)hipFree(b2);
  rc = build_graphs(s);
  if (rc) return rc;
  if (chosen) {
    if (s->flat2d) ens_tile_cells(bt, chosen[0], chosen[1]);
    else fused_tile_cells(bt, chosen[0], chosen[1]);
    chosen[2] = s->zc;
  }
  return SHX_OK;
}

}  // extern "C"
