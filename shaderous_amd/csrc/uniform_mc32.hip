// FP32 uniform-grid kernels with the monotonized-central limiter (PLM).
#include "uniform_impl.hpp"
namespace shx {
UniformOps uniform_ops_mc32(int rs, int isen) {
  if (rs == RS_HLLC) return isen ? UniformImpl<float, LIM_MC, RS_HLLC, true>::ops() : UniformImpl<float, LIM_MC, RS_HLLC, false>::ops();
  return isen ? UniformImpl<float, LIM_MC, RS_HLL, true>::ops() : UniformImpl<float, LIM_MC, RS_HLL, false>::ops();
}
}  // namespace shx
