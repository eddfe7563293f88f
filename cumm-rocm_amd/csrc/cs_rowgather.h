// Row-gather implicit GEMM kernel (sparse conv forward / dgrad, ShuffleAC and
// dense GEMM).  One wave owns 64 sorted rows (four 16-row mask tiles) and
// BN = 16*NB output columns.  Gathered rows and weight rows are loaded straight
// into MFMA operand registers (16 B per lane); no LDS and no block barrier.
// The wave ORs its four tile masks into a scalar union and walks it with
// s_ff1; per offset, tiles whose own mask lacks the bit are skipped with a
// scalar branch.  Instantiated per input dtype in rg_<dtype>.hip.
#pragma once
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

template <typename T, typename OutT, int NB, int KCH>
__global__ __launch_bounds__(256) void rowgather_kernel(RowGatherArgs p) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int row0 = (blockIdx.x * kWavesPerBlock + wave) * kRowsPerWave;
  if (row0 >= p.m) return;  // whole wave leaves together
  const int n0 = blockIdx.y * (NB * 16);
  const int lr = lane & 15;
  const int lg = lane >> 4;

  const int myrow = row0 + lane;
  const bool rvalid = myrow < p.m;
  const uint32_t kvmask = p.kv >= 32 ? 0xffffffffu : ((1u << p.kv) - 1u);
  uint32_t mymask = 0;
  int mycol = 0, myout = 0;
  if (rvalid) {
    mymask = p.mask ? p.mask[myrow] : kvmask;
    mycol = p.pair_col ? p.pair_col[myrow] : myrow;
    myout = p.out_row ? p.out_row[myrow] : myrow;
  }
  // OR of the 16 row masks of each tile (lanes 16s..16s+15 hold tile s).
  uint32_t g = mymask;
  g |= __shfl_xor(g, 1);
  g |= __shfl_xor(g, 2);
  g |= __shfl_xor(g, 4);
  g |= __shfl_xor(g, 8);
  uint32_t sm[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) sm[s] = __builtin_amdgcn_readlane(g, 16 * s) & p.mask_filter & kvmask;

  if (p.mask_out != nullptr && lane < 4 && row0 + 16 * lane < p.m) {
    const uint32_t v = lane == 0 ? sm[0] : lane == 1 ? sm[1] : lane == 2 ? sm[2] : sm[3];
    p.mask_out[row0 / kMaskTile + lane] = v;
  }
  uint32_t live = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (row0 + 16 * s < p.m && (!p.skip_empty || sm[s] != 0u)) live |= 1u << s;
  }
  if (p.reverse_mask) {
#pragma unroll
    for (int s = 0; s < 4; ++s) sm[s] = __builtin_bitreverse32(sm[s]) >> (32 - p.kv);
  }
  uint32_t uni = (sm[0] | sm[1] | sm[2] | sm[3]);

  int cols[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) cols[s] = __shfl(mycol, 16 * s + lr);

  f32x4 acc[4][NB];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[s][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  const T* A = reinterpret_cast<const T*>(p.a);
  while (uni) {
    const int f = __builtin_ctz(uni);
    uni &= uni - 1u;
    int idx[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      idx[s] = -1;
      if ((sm[s] >> f) & 1u) idx[s] = p.pairs ? p.pairs[(int64_t)f * p.pair_ld + cols[s]] : cols[s];
    }
    const T* bf = reinterpret_cast<const T*>(p.b) + (int64_t)f * p.b_fs;
    for (int kb = 0; kb < p.kred; kb += KCH * CH) {
#pragma unroll
      for (int kc = 0; kc < KCH; ++kc) {
        const int k = kb + kc * CH + lg * E;
        const bool kin = k < p.kred;
        u32x4 bfr[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          const int n = n0 + nb * 16 + lr;
          bfr[nb] = (kin && n < p.nout) ? ld16(bf + (int64_t)n * p.b_ns + k) : zero16();
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if ((sm[s] >> f) & 1u) {
            const u32x4 afr = (kin && idx[s] >= 0) ? ld16(A + (int64_t)idx[s] * p.lda + k) : zero16();
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) MM::step(acc[s][nb], afr, bfr[nb]);
          }
        }
      }
    }
  }

  OutT* C = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (!((live >> s) & 1u)) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = 16 * s + 4 * lg + i;
      const int orow = __shfl(myout, rr);
      if (row0 + rr < p.m) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          const int n = n0 + nb * 16 + lr;
          if (n < p.nout) {
            float v = acc[s][nb][i];
            if (p.scale) v *= p.scale[n];
            v *= p.alpha;
            if (p.beta != 0.f) {
              const float src = p.d_is_bias ? to_f32(D[n]) : to_f32(D[(int64_t)orow * p.ldd + n]);
              v += p.beta * src;
            }
            v = apply_act(v, p.act_type, p.act_alpha);
            C[(int64_t)orow * p.ldc + n] = from_f32<OutT>(v);
          }
        }
      }
    }
  }
}


template <typename T, typename OutT, int NB>
static hipError_t launch_rg_nb(const RowGatherArgs& a, hipStream_t s) {
  constexpr int CH = Mma<T>::kChunk;
  const int chunks = (a.kred + CH - 1) / CH;
  dim3 grid((a.m + kRowsPerWave * kWavesPerBlock - 1) / (kRowsPerWave * kWavesPerBlock),
            (a.nout + NB * 16 - 1) / (NB * 16), 1);
  dim3 block(64 * kWavesPerBlock, 1, 1);
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  if (chunks <= 1)
    hipLaunchKernelGGL((rowgather_kernel<T, OutT, NB, 1>), grid, block, 0, s, a);
  else if (chunks <= 2)
    hipLaunchKernelGGL((rowgather_kernel<T, OutT, NB, 2>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((rowgather_kernel<T, OutT, NB, 4>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <typename T, typename OutT>
static hipError_t launch_rg(const RowGatherArgs& a, int block_n, hipStream_t s) {
  if (block_n == 64) return launch_rg_nb<T, OutT, 4>(a, s);
  if (block_n == 128) return launch_rg_nb<T, OutT, 8>(a, s);
  return hipErrorInvalidValue;
}


}  // namespace cs
