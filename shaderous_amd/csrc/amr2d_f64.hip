// FP64 block-AMR kernels for 2D grids: the all-FP64 twin of the production path (precision
// study), MC and Koren with the density reconstruction along the adiabat, HLLC.
#include "amr_impl.hpp"
namespace shx {
AmrOps amr2d_ops_f64(int lim, int rs, int isen) {
  if (rs != RS_HLLC) return AmrOps{};
  if (lim == LIM_MC && !isen) return AmrImpl<double, LIM_MC, RS_HLLC, false, 2>::ops();
  if (lim == LIM_KOREN && isen) return AmrImpl<double, LIM_KOREN, RS_HLLC, true, 2>::ops();
  return AmrOps{};
}
}  // namespace shx
