// Gather GEMM kernel instantiations for f32 operands.
#include "cs_gemm.h"

namespace cs {

hipError_t launch_gemm_rows_f32(int out_dtype, const GemmRowsArgs& a, hipStream_t s) {
  if (out_dtype == kF32) return launch_gm<float, float>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
