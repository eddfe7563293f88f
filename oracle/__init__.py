"""CPU oracle for the gather-GEMM / mask-sparse implicit-GEMM path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``cumm-rocm_amd/``) imports
this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or the
timed CPU baseline.

Contents
  spconv_ref.py    FP64 numpy restatement of implicit_gemm2's forward / dgrad /
                   wgrad semantics (tile masks, filter, reverse, epilogue) and the
                   reference's CPU gather-GEMM-scatter path over compact pairs.
  gemm_ref.py      FP64 restatement of matmul2 (NoShuffle / ShuffleAC /
                   ShuffleAB) and the reference's blocked FP32 ``py_matmul``.
  bruteforce.py    pure-Python neighbour search + accumulation straight from
                   voxel coordinates (tiny inputs), independent of pair tables.
  gemm_emu.c       C emulation of the FP32 row-gather kernel's accumulation
                   order (stand-in for the reference's cudasim GEMM run).
  ref_host.cpp     C-ABI shim compiled against the reference's own header-only
                   host code (params.h / utils.h) into _ref/, when the reference
                   tree is present.
  Makefile         builds _build/libgemm_emu.so and _ref/libcumm_ref_host.so.

How each function is pinned is stated in its docstring; the tests in
``tests/test_oracle.py`` carry the checks.
"""
