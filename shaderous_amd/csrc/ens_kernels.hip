// Ensemble (batched 2D) kernels: FP32 with MC (HLLC or HLL) or Koren (+- isentropic) and HLLC,
// the other FP32 schemes in ens_other32.hip; the FP64 twin with MC or Koren + isentropic, HLLC.
#include "ens_impl.hpp"
namespace shx {
EnsOps ens_ops_f32(int lim, int rs, int isen) {
  if (lim == LIM_MC && !isen) return rs == RS_HLLC ? EnsImpl<float, LIM_MC, RS_HLLC, false>::ops() : EnsImpl<float, LIM_MC, RS_HLL, false>::ops();
  if (lim == LIM_KOREN && rs == RS_HLLC) return isen ? EnsImpl<float, LIM_KOREN, RS_HLLC, true>::ops() : EnsImpl<float, LIM_KOREN, RS_HLLC, false>::ops();
  return ens_ops_other32(lim, rs, isen);
}
EnsOps ens_ops_f64(int lim, int rs, int isen) {
  if (rs != RS_HLLC) return EnsOps{};
  if (lim == LIM_MC && !isen) return EnsImpl<double, LIM_MC, RS_HLLC, false>::ops();
  if (lim == LIM_KOREN && isen) return EnsImpl<double, LIM_KOREN, RS_HLLC, true>::ops();
  return EnsOps{};
}
}  // namespace shx
