// Row-gather kernel instantiations for e4m3 inputs.
#include "cs_rowgather.h"

namespace cs {

hipError_t launch_rowgather_e4m3(int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  if (out_dtype == kF16) return launch_rg<e4m3, _Float16>(a, block_n, s);
  if (out_dtype == kBF16) return launch_rg<e4m3, bf16>(a, block_n, s);
  if (out_dtype == kF32) return launch_rg<e4m3, float>(a, block_n, s);
  return hipErrorInvalidValue;
}

}  // namespace cs
