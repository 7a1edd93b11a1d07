// FP32 block-AMR kernels: PLM (MC) and the reference defaults (Koren, with or without the
// density reconstruction along the adiabat), HLLC or HLL.
#include "amr_fused.hpp"
namespace shx {
AmrOps amr_ops_f32(int lim, int rs, int isen) {
  if (lim == LIM_MC && !isen) return rs == RS_HLLC ? AmrImpl<float, LIM_MC, RS_HLLC, false>::ops() : AmrImpl<float, LIM_MC, RS_HLL, false>::ops();
  if (lim == LIM_KOREN && rs == RS_HLLC) return isen ? AmrImpl<float, LIM_KOREN, RS_HLLC, true>::ops() : AmrImpl<float, LIM_KOREN, RS_HLLC, false>::ops();
  return AmrOps{};
}
}  // namespace shx
