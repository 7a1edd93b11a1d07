// Ensemble (batched 2D) kernels in FP32 with the van Leer and minmod limiters (HLLC or HLL) and
// with Koren and HLL (with or without the density reconstruction along the adiabat).
#include "ens_impl.hpp"
namespace shx {
EnsOps ens_ops_other32(int lim, int rs, int isen) {
  if (lim == LIM_VANLEER && !isen) return rs == RS_HLLC ? EnsImpl<float, LIM_VANLEER, RS_HLLC, false>::ops() : EnsImpl<float, LIM_VANLEER, RS_HLL, false>::ops();
  if (lim == LIM_MINMOD && !isen) return rs == RS_HLLC ? EnsImpl<float, LIM_MINMOD, RS_HLLC, false>::ops() : EnsImpl<float, LIM_MINMOD, RS_HLL, false>::ops();
  if (lim == LIM_KOREN && rs == RS_HLL) return isen ? EnsImpl<float, LIM_KOREN, RS_HLL, true>::ops() : EnsImpl<float, LIM_KOREN, RS_HLL, false>::ops();
  return EnsOps{};
}
}  // namespace shx
