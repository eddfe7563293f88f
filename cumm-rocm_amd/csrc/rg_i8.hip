// Row-gather kernel instantiations for int8 inference (int8 in, int8 out,
// int32 accumulation).
#include "cs_rowgather.h"

namespace cs {

hipError_t launch_rowgather_i8(int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  if (out_dtype == kI8) return launch_rg<int8_t, int8_t>(a, block_n, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
