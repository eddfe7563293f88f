"""cumm_amd: MI355X (gfx950) implementation of cumm's gather-GEMM and
mask-sparse implicit-GEMM sparse-convolution path.

Public surface (mirrors cumm):
  cumm_amd.tensorview        -- ``tv`` (Tensor, dtypes, tv.gemm params/desps)
  cumm_amd.conv.main         -- ConvMainUnitTest.implicit_gemm2
  cumm_amd.gemm.main         -- GemmMainUnitTest.matmul2
"""
__version__ = "0.1.0"
