// FP32 block-AMR kernels with the van Leer and minmod limiters (HLLC or HLL).
#include "amr_fused.hpp"
namespace shx {
AmrOps amr_ops_other32(int lim, int rs) {
  if (lim == LIM_VANLEER) return rs == RS_HLLC ? AmrImpl<float, LIM_VANLEER, RS_HLLC, false>::ops() : AmrImpl<float, LIM_VANLEER, RS_HLL, false>::ops();
  if (lim == LIM_MINMOD) return rs == RS_HLLC ? AmrImpl<float, LIM_MINMOD, RS_HLLC, false>::ops() : AmrImpl<float, LIM_MINMOD, RS_HLL, false>::ops();
  return AmrOps{};
}
}  // namespace shx
