// FP32 block-AMR kernels with the Koren limiter and HLL (with or without the density
// reconstruction along the adiabat).
#include "amr_fused.hpp"
namespace shx {
AmrOps amr_ops_koren_hll32(int isen) {
  return isen ? AmrImpl<float, LIM_KOREN, RS_HLL, true>::ops() : AmrImpl<float, LIM_KOREN, RS_HLL, false>::ops();
}
}  // namespace shx
