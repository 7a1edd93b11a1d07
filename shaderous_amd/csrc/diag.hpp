// The diagnostics reduction (hydro/diagnostics.wgsl:13-98 on the uniform grid,
// amr/diagnostics.wgsl:13-117 over leaves): 24 channels per cell, volume-weighted, summed in
// FP64 in a fixed order (deterministic). Channel layout and operations are the reference's
// (DIAG_OPS, src/sim.ts:1808):
//   [0] mass, [1..3] momentum, [4] kinetic energy, [5] internal energy as the total holds it
//   (E - KE), [6] gravitational energy (none on this path: 0), [7] max Mach, [8] floored cells,
//   [9] Jeans-unresolved cells (0), [10, 11] rho min / max, [12, 13] T min / max,
//   [14, 15] s min / max, [16, 17] p min / max, [18, 19] speed min / max,
//   [20..22] rho dv x (the centre of mass times the mass), [23] nuclear heating (0).
#pragma once
#include "hydro.hpp"

namespace shx {

constexpr int DIAG_CH = 24;
constexpr double DIAG_BIG = 1e300;

// 0 sum, 1 max, 2 min (DIAG_OPS)
__host__ __device__ constexpr int diag_op(int c) {
  return c == 7 ? 1 : (c >= 10 && c <= 19) ? ((c & 1) ? 1 : 2) : 0;
}
__device__ __forceinline__ double diag_fold(int c, double a, double b) {
  const int op = diag_op(c);
  return op == 0 ? a + b : op == 1 ? fmax(a, b) : fmin(a, b);
}
__device__ __forceinline__ double diag_id(int c) {
  const int op = diag_op(c);
  return op == 0 ? 0.0 : op == 1 ? -DIAG_BIG : DIAG_BIG;
}

// Where the cells are: dv the volume of one cell (dx^ndim: a 2D run's quantities are per unit
// length along z, sim.ts:1733), org the domain's lower corner (Params.org, src/domain.ts:1-40).
struct DiagGeo {
  double dv;
  double org[3];
  int flat;            // a 2D run lies in the plane z = 0 (diagnostics.wgsl:55)
  // the gas and floors as configured (not rounded to the run's precision)
  double gamma, rho_floor, p_floor;
};

// red_cell (hydro/diagnostics.wgsl:13-98) for one cell at (x, y, z), ideal gas, no gravity,
// burning or partial cells, accumulated into r. A solid cell (mark > 1/2) holds no gas.
template <typename T>
__device__ __forceinline__ void diag_cell(const DiagGeo& G, T4<T> u1, T4<T> u2, double dv, double x, double y,
                                          double z, double r[DIAG_CH]) {
  if ((double)u2.w > 0.5) return;
  const double gamma = G.gamma, gm1 = gamma - 1.0;
  const double rho_floor = G.rho_floor, p_floor = G.p_floor;
  const double rho = fmax((double)u1.x, rho_floor);
  const double vx = (double)u1.y / rho, vy = (double)u1.z / rho, vz = (double)u1.w / rho;
  const double eint = (double)u2.y;
  const double p = fmax(gm1 * eint, p_floor);
  const double cs = sqrt(gamma * p / rho);
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double speed = sqrt(v2);
  const double temp = gm1 * eint / rho;
  const double s = (double)u2.z / rho;
  const double m = rho * dv;
  r[0] += m;
  r[1] += (double)u1.y * dv;
  r[2] += (double)u1.z * dv;
  r[3] += (double)u1.w * dv;
  r[4] += 0.5 * rho * v2 * dv;
  r[5] += ((double)u2.x - 0.5 * rho * v2) * dv;
  r[7] = fmax(r[7], speed / cs);
  r[8] += ((double)u1.x <= rho_floor * 1.0001 || p <= p_floor * 1.0001) ? 1.0 : 0.0;
  r[10] = fmin(r[10], rho); r[11] = fmax(r[11], rho);
  r[12] = fmin(r[12], temp); r[13] = fmax(r[13], temp);
  r[14] = fmin(r[14], s); r[15] = fmax(r[15], s);
  r[16] = fmin(r[16], p); r[17] = fmax(r[17], p);
  r[18] = fmin(r[18], speed); r[19] = fmax(r[19], speed);
  r[20] += m * x;
  r[21] += m * y;
  r[22] += m * z;
}

__device__ __forceinline__ void diag_init(double r[DIAG_CH]) {
#pragma unroll
  for (int c = 0; c < DIAG_CH; c++) r[c] = diag_id(c);
}

// Fold a workgroup's per-thread records into out[DIAG_CH]: a fixed butterfly inside each wave,
// then the waves in order (deterministic). sh holds (blockDim.x / 64) x DIAG_CH doubles.
__device__ __forceinline__ void diag_block(double r[DIAG_CH], double* sh, double* out) {
#pragma unroll
  for (int c = 0; c < DIAG_CH; c++)
    for (int m = 32; m > 0; m >>= 1) r[c] = diag_fold(c, r[c], __shfl_xor(r[c], m, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
    for (int c = 0; c < DIAG_CH; c++) sh[w * DIAG_CH + c] = r[c];
  __syncthreads();
  if ((int)threadIdx.x < DIAG_CH) {
    const int c = threadIdx.x;
    double a = sh[c];
    for (int k = 1; k < nw; k++) a = diag_fold(c, a, sh[k * DIAG_CH + c]);
    out[c] = a;
  }
}

// The uniform grid: one record of the channels per workgroup over a fixed set of cells (a
// grid-stride loop), folded by k_diag_final in workgroup order.
template <typename T>
__global__ __launch_bounds__(256) void k_diag(Params<T> P, DiagGeo G, const T4<T>* __restrict__ U1,
                                              const T4<T>* __restrict__ U2, int ncell, double* part) {
  __shared__ double sh[4 * DIAG_CH];
  double r[DIAG_CH];
  diag_init(r);
  const int nx = P.n[0], ny = P.n[1];
  const double dx = (double)P.dx;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < ncell; o += gridDim.x * blockDim.x) {
    const int i = o % nx, j = (o / nx) % ny, k = o / (nx * ny);
    const double x = (i + 0.5) * dx + G.org[0], y = (j + 0.5) * dx + G.org[1];
    const double z = G.flat ? 0.0 : (k + 0.5) * dx + G.org[2];
    diag_cell<T>(G, U1[o], U2[o], G.dv, x, y, z, r);
  }
  diag_block(r, sh, part + (size_t)blockIdx.x * DIAG_CH);
}

// The channels over `nparts` partial records: 32 groups of consecutive records per channel,
// each folded in record order, then the groups in order (a fixed order: deterministic).
// Launch: one workgroup of 32 x DIAG_CH threads. (T: the state's precision, which the records
// do not depend on; a template, so every translation unit may instantiate it.)
template <typename T>
__global__ __launch_bounds__(32 * DIAG_CH) void k_diag_final(const double* part, int nparts, double* out) {
  __shared__ double grp[32][DIAG_CH];
  const int c = threadIdx.x % DIAG_CH, g = threadIdx.x / DIAG_CH;
  const int per = (nparts + 31) / 32, b0 = g * per, b1 = min(b0 + per, nparts);
  double a = diag_id(c);
  for (int b = b0; b < b1; b++) a = diag_fold(c, a, part[(size_t)b * DIAG_CH + c]);
  grp[g][c] = a;
  __syncthreads();
  if (g == 0) {
    double v = diag_id(c);
    for (int k = 0; k < 32; k++) v = diag_fold(c, v, grp[k][c]);
    out[c] = v;
  }
}

}  // namespace shx
