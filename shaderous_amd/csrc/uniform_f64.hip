// FP64 uniform-grid kernels: the all-FP64 twin of the production path (precision study).
#include "uniform_impl.hpp"
namespace shx {
UniformOps uniform_ops_f64(int lim, int rs, int isen) {
  if (rs != RS_HLLC) return UniformOps{};
  if (lim == LIM_MC && !isen) return UniformImpl<double, LIM_MC, RS_HLLC, false>::ops();
  if (lim == LIM_KOREN) return isen ? UniformImpl<double, LIM_KOREN, RS_HLLC, true>::ops() : UniformImpl<double, LIM_KOREN, RS_HLLC, false>::ops();
  return UniformOps{};
}
}  // namespace shx
