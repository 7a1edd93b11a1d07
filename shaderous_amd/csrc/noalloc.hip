// The diagnostics reduction over a caller's device state with no allocation (include/shx.h,
// shx_diagnostics_noalloc): for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops). The per-workgroup records that shx_diagnostics keeps in the sim's
// own buffer live here in a file-scope device array, `scratch`, one copy per GPU, held until the
// process ends. Host code never passes the array's name to a kernel: on each device its address
// is taken once with hipGetSymbolAddress (with that device current) and kept per device.
#include <hip/hip_runtime.h>
#include <limits.h>

#include <mutex>

#include "common.hpp"
#include "diag.hpp"

namespace shx {
namespace {

constexpr int NOALLOC_RECORDS = 256;    // one record per workgroup of k_diag (a grid-stride loop)
constexpr int NOALLOC_MAX_DEVICES = 64;

// 256 records x 24 channels x 8 B = 48 KiB per GPU
__device__ __align__(256) double scratch[NOALLOC_RECORDS * DIAG_CH];

std::mutex g_scratch_mu;
double* g_scratch[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch` (its address on that device, taken on first use).
int scratch_of_current_device(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(SHX_EDEVICE, "device %d beyond the %d the library tracks", dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(SHX_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return SHX_OK;
}

template <typename T>
int launch(const shx_config& c, const void* u1, const void* u2, double* out, hipStream_t st, double* ptr) {
  const Params<T> P = make_params<T>(c);
  const int ncell = c.n[0] * c.n[1] * c.n[2];
  hipLaunchKernelGGL(k_diag<T>, dim3(NOALLOC_RECORDS), dim3(256), 0, st, P, uniform_diag_geo(c), (const T4<T>*)u1,
                     (const T4<T>*)u2, ncell, ptr);
  hipLaunchKernelGGL(k_diag_final<T>, dim3(1), dim3(32 * DIAG_CH), 0, st, (const double*)ptr, NOALLOC_RECORDS, out);
  HIPCHK(hipGetLastError());
  return SHX_OK;
}

}  // namespace
}  // namespace shx

using namespace shx;

extern "C" int shx_diagnostics_noalloc(const shx_config* cfg, const void* u1, const void* u2, double* out,
                                       void* stream) {
  if (!cfg || !u1 || !u2 || !out) return fail(SHX_EINVAL, "null argument");
  const shx_config& c = *cfg;
  for (int a = 0; a < 3; a++)
    if (c.n[a] < 1) return fail(SHX_EINVAL, "n[%d] = %d", a, c.n[a]);
  if (c.precision != 32 && c.precision != 64) return fail(SHX_EINVAL, "precision %d", c.precision);
  if (!(c.gamma > 1) || !(c.dx > 0)) return fail(SHX_EINVAL, "gamma must exceed 1 and dx be positive");
  const double cells = (double)c.n[0] * c.n[1] * c.n[2];
  if (cells > (double)INT_MAX)
    return fail(SHX_EINVAL, "a grid of %.0f cells: the no-allocation reduction covers at most %d cells", cells, INT_MAX);
  double* ptr = nullptr;
  int rc = scratch_of_current_device(&ptr);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return c.precision == 64 ? launch<double>(c, u1, u2, out, st, ptr) : launch<float>(c, u1, u2, out, st, ptr);
}
