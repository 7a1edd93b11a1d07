// Shared host-side helpers of libshx: the per-thread error message behind shx_last_error(), the
// HIP status check, and the device parameter block built from a shx_config.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <string>

#include "../../include/shx.h"
#include "diag.hpp"
#include "hydro.hpp"

namespace shx {
extern thread_local std::string g_err;
// Record a formatted message and return `code` (a negative SHX_E* status).
int fail(int code, const char* fmt, ...);

// The Params block of a configuration, in precision T.
template <typename T>
Params<T> make_params(const shx_config& c) {
  Params<T> P;
  for (int a = 0; a < 3; a++) { P.n[a] = c.n[a]; P.bc_lo[a] = c.bc_lo[a]; P.bc_hi[a] = c.bc_hi[a]; }
  P.flat = c.n[2] == 1;
  P.ndim = T(c.ndim);
  P.gamma = T(c.gamma);
  P.gm1 = P.gamma - T(1);
  P.igm1 = T(1) / P.gm1;
  P.dx = T(c.dx);
  P.cfl = T(c.cfl);
  P.rho_floor = T(c.rho_floor);
  P.p_floor = T(c.p_floor);
  P.eta_dual = T(c.eta_dual);
  P.shock_thresh = T(c.shock_thresh);
  double te = c.t_end;
  if (!(te < 1e30)) te = 1e30;
  P.t_end = T(te);
  P.dt_max = T(1e30);
  for (int a = 0; a < 3; a++) P.fluff[a] = T(c.fluff[a]);
  if (!(c.fluff[0] > 0)) { P.fluff[0] = T(0); P.fluff[1] = T(0); P.fluff[2] = T(0); }
  else {
    if (!(c.fluff[1] > 0)) P.fluff[1] = T(1e30);
    if (!(c.fluff[2] > 0)) P.fluff[2] = T(1e30);
  }
  return P;
}

// Where a uniform grid's cells are, for the diagnostics: dv = dx^ndim, the lower corner.
inline DiagGeo uniform_diag_geo(const shx_config& c) {
  DiagGeo G;
  G.dv = c.ndim == 1 ? c.dx : c.ndim == 2 ? c.dx * c.dx : c.dx * c.dx * c.dx;
  for (int a = 0; a < 3; a++) G.org[a] = c.org[a];
  G.flat = c.n[2] == 1;
  G.gamma = c.gamma;
  G.rho_floor = c.rho_floor;
  G.p_floor = c.p_floor;
  return G;
}

// Sim.runToEnd's loop (src/sim.ts:3689-3718) over a sim's step and clock read: batches double
// from `batch` up to 512 but never run far past the steps the clock says remain; done when the
// clock holds (dt = 0, t > 0), with *steps = its banked count. NaN clocks and passing max_steps
// before the end are errors, as the reference throws for both.
template <typename Step, typename Clock>
int run_to_end_loop(Step step, Clock read_clock, double t_end, int batch, int max_steps, int* steps) {
  if (batch < 1) batch = 16;
  int size = batch;
  double c[11];
  for (;;) {
    int rc = step(size);
    if (rc) return rc;
    rc = read_clock(c);
    if (rc) return rc;
    const double dt = c[0], t = c[1];
    if (!(dt == dt) || !(t == t)) return fail(SHX_ESTATE, "clock is NaN after %d steps", (int)c[3]);
    const int banked = (int)c[3];
    if (dt == 0 && t > 0) {
      if (steps) *steps = banked;
      return SHX_OK;
    }
    if (banked > max_steps) return fail(SHX_ESTATE, "no convergence to t_end after %d steps (t = %g)", banked, t);
    size = size * 2 < 512 ? size * 2 : 512;
    if (dt > 0) {
      const double left = (t_end - t - dt) / dt * 1.1 + 1.0;
      if (left < size) size = left < 1.0 ? 1 : (int)ceil(left);
    }
  }
}

// An asynchronous readback (include/shx.h): copies enqueued on a sim's stream into pinned host
// memory it owns, and an event recorded after them.
struct Readback {
  hipEvent_t ev = nullptr;
  unsigned char* host = nullptr;   // [u1 | u2 | clock (16 values) | diagnostics (24 doubles)]
  size_t ncell = 0;
  int es = 4;
  int what = 0;
  bool done = false;
  double clock[11] = {};
  double diag[24] = {};
};
// Enqueue the copies of a readback on `stream`: the state U1, U2 (ncell cells of 4 values of es
// bytes) if what & 1, the 16 clock values, and the DIAG_CH doubles at diag (already computed on
// the stream) if what & 2.
int readback_begin(hipStream_t stream, int what, int es, size_t ncell, const void* U1, const void* U2, const void* clk,
                   const double* diag, shx_readback** out);

}  // namespace shx

struct shx_readback : shx::Readback {};

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(SHX_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)
