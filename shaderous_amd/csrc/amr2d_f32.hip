// FP32 block-AMR kernels for 2D grids (quadtrees of 16^2-cell blocks), the passes forms (plain
// per level and batched): PLM with MC, van Leer or minmod (HLLC or HLL) and the reference
// defaults (Koren with HLLC or HLL, with or without the density reconstruction along the adiabat).
#include "amr_impl.hpp"
namespace shx {
AmrOps amr2d_ops_f32(int lim, int rs, int isen) {
  const bool c = rs == RS_HLLC;
  if (lim == LIM_KOREN) {
    if (isen) return c ? AmrImpl<float, LIM_KOREN, RS_HLLC, true, 2>::ops() : AmrImpl<float, LIM_KOREN, RS_HLL, true, 2>::ops();
    return c ? AmrImpl<float, LIM_KOREN, RS_HLLC, false, 2>::ops() : AmrImpl<float, LIM_KOREN, RS_HLL, false, 2>::ops();
  }
  if (isen) return AmrOps{};
  if (lim == LIM_MC) return c ? AmrImpl<float, LIM_MC, RS_HLLC, false, 2>::ops() : AmrImpl<float, LIM_MC, RS_HLL, false, 2>::ops();
  if (lim == LIM_VANLEER) return c ? AmrImpl<float, LIM_VANLEER, RS_HLLC, false, 2>::ops() : AmrImpl<float, LIM_VANLEER, RS_HLL, false, 2>::ops();
  if (lim == LIM_MINMOD) return c ? AmrImpl<float, LIM_MINMOD, RS_HLLC, false, 2>::ops() : AmrImpl<float, LIM_MINMOD, RS_HLL, false, 2>::ops();
  return AmrOps{};
}
}  // namespace shx
