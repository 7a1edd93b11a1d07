// FP32 uniform-grid kernels with the Koren limiter (the reference's default).
#include "uniform_impl.hpp"
namespace shx {
UniformOps uniform_ops_koren32(int rs, int isen) {
  if (rs == RS_HLLC) return isen ? UniformImpl<float, LIM_KOREN, RS_HLLC, true>::ops() : UniformImpl<float, LIM_KOREN, RS_HLLC, false>::ops();
  return isen ? UniformImpl<float, LIM_KOREN, RS_HLL, true>::ops() : UniformImpl<float, LIM_KOREN, RS_HLL, false>::ops();
}
}  // namespace shx
