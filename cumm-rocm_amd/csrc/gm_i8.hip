// Gather GEMM kernel instantiations for i8 operands.
#include "cs_gemm.h"

namespace cs {

hipError_t launch_gemm_rows_i8(int out_dtype, const GemmRowsArgs& a, hipStream_t s) {
  if (out_dtype == kI8) return launch_gm<int8_t, int8_t>(a, s);
  if (out_dtype == kI32) return launch_gm<int8_t, int32_t>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
