// FP32 uniform-grid kernels with the van Leer and minmod limiters.
#include "uniform_impl.hpp"
namespace shx {
UniformOps uniform_ops_other32(int lim, int rs) {
  if (lim == LIM_VANLEER) return rs == RS_HLLC ? UniformImpl<float, LIM_VANLEER, RS_HLLC, false>::ops() : UniformImpl<float, LIM_VANLEER, RS_HLL, false>::ops();
  if (lim == LIM_MINMOD) return rs == RS_HLLC ? UniformImpl<float, LIM_MINMOD, RS_HLLC, false>::ops() : UniformImpl<float, LIM_MINMOD, RS_HLL, false>::ops();
  return UniformOps{};
}
}  // namespace shx
