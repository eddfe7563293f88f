/*
 * CPU emulation of the FP32 row-gather GEMM kernel's accumulation order.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Stand-in for the
 * reference's cudasim run of its FP32 GEMM kernel (cumm/gemm/main.py:1559-1660,
 * cumm/cudasim/__init__.py:236-283), which needs pccm and cannot run here.
 *
 * For C[i][j] = sum_k A[i][k] * Bt[j][k] (A row-major [m][k], Bt row-major [n][k])
 * the FP32 kernel walks k in chunks of 16; in chunk c, lane group g (0..3)
 * holds elements 16c + 4g + q (q = 0..3) and step q issues one 16x16x4 MFMA
 * that adds the four products {16c + 4g + q : g = 0..3} to the FP32
 * accumulator.  The emulation sums those four products in g order, in FP32,
 * then adds them to the accumulator.  Elements past k are zero.
 * Pinned against numpy FP64 within FP32 rounding in tests/test_oracle.py.
 */
#include <stdint.h>

void gemm_emu_f32(const float* A, const float* Bt, float* C, int64_t m, int64_t n, int64_t k) {
  for (int64_t i = 0; i < m; ++i) {
    for (int64_t j = 0; j < n; ++j) {
      float acc = 0.0f;
      for (int64_t c = 0; c < k; c += 16) {
        for (int q = 0; q < 4; ++q) {
          float s = 0.0f;
          for (int g = 0; g < 4; ++g) {
            const int64_t kk = c + 4 * g + q;
            if (kk < k) s += A[i * k + kk] * Bt[j * k + kk];
          }
          acc += s;
        }
      }
      C[i * n + j] = acc;
    }
  }
}
