// Gather GEMM kernel instantiations for bf16 operands.
#include "cs_gemm.h"

namespace cs {

hipError_t launch_gemm_rows_bf16(int out_dtype, const GemmRowsArgs& a, hipStream_t s) {
  if (out_dtype == kBF16) return launch_gm<bf16, bf16>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
