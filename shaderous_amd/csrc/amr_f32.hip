// FP32 block-AMR kernels: PLM (MC, HLLC or HLL) and the reference defaults (Koren with HLLC, with
// or without the density reconstruction along the adiabat); the other limiters and Koren with HLL
// in amr_other32.hip and amr_koren_hll32.hip.
#include "amr_fused.hpp"
namespace shx {
AmrOps amr_ops_f32(int lim, int rs, int isen) {
  if (lim == LIM_MC && !isen) return rs == RS_HLLC ? AmrImpl<float, LIM_MC, RS_HLLC, false>::ops() : AmrImpl<float, LIM_MC, RS_HLL, false>::ops();
  if (lim == LIM_KOREN && rs == RS_HLLC) return isen ? AmrImpl<float, LIM_KOREN, RS_HLLC, true>::ops() : AmrImpl<float, LIM_KOREN, RS_HLLC, false>::ops();
  if (lim == LIM_KOREN) return amr_ops_koren_hll32(isen);
  if (!isen) return amr_ops_other32(lim, rs);
  return AmrOps{};
}
}  // namespace shx
