// Shared host-side helpers of libshx: the per-thread error message behind shx_last_error(), the
// HIP status check, and the device parameter block built from a shx_config.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "../../include/shx.h"
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
}  // namespace shx

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(SHX_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)
