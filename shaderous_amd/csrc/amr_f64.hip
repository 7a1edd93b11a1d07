// FP64 block-AMR kernels: the all-FP64 twin of the production path (precision study).
#include "amr_fused.hpp"
namespace shx {
AmrOps amr_ops_f64(int lim, int rs, int isen) {
  if (rs != RS_HLLC) return AmrOps{};
  if (lim == LIM_MC && !isen) return AmrImpl<double, LIM_MC, RS_HLLC, false>::ops();
  if (lim == LIM_KOREN && isen) return AmrImpl<double, LIM_KOREN, RS_HLLC, true>::ops();
  return AmrOps{};
}
}  // namespace shx
