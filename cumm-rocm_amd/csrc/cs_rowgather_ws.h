// Weight-stationary row-gather kernel (sparse conv forward / dgrad).
//
// Measured on MI355X, a dependent gather round trip under full load costs
// ~1.5-2 us, and a kernel whose waves walk a tile's kernel offsets one round
// trip at a time lasts as long as its heaviest tiles (up to 24 offsets).  This
// kernel bounds every tile to three round trips (masks -> pair indices of up
// to 8 offsets -> their gathered rows) and keeps register use low so 16 waves
// fit on a CU:
//   * persistent blocks, one per CU: the weights of all kv offsets for a slice
//     of NCOL output columns are loaded once into LDS (<= 108 KiB, 16-byte
//     chunks XOR-swizzled by row so MFMA operand reads are conflict-free);
//   * each wave takes 16-row mask tiles round-robin over the whole grid;
//   * per tile: one load round trip for row masks / argsort, one for the pair
//     indices of up to 8 offsets (each of the four lane groups loads a
//     different offset), one for the gathered rows of those offsets (raw buffer
//     loads, index -1 reads zero); MFMAs take weights from LDS.
// Used when kv * NCOL * C * sizeof(T) fits the LDS budget with NCOL >= 16.
#pragma once
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

constexpr int kWsLdsBytes = 110592;  // 108 KiB of weights per block
constexpr int kWsWaves = 16;         // waves per block

// KC: reduction chunks (C / chunk); NBC: 16-column blocks per wave (NCOL / 16)
template <typename T, typename OutT, int KC, int NBC>
__global__ __launch_bounds__(1024) void rowgather_ws_kernel(RowGatherArgs p, int ntiles) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int NCOL = NBC * 16;
  constexpr int C = KC * CH;            // reduction length handled (== p.kred)
  constexpr int CHUNKS = C / E;         // 16-byte chunks per weight row
  constexpr int SWZ = CHUNKS >= 8 ? 7 : CHUNKS - 1;
  constexpr int BATCH = KC > 2 ? 4 : 8;  // offsets per gather round trip
  __shared__ __attribute__((aligned(16))) char wlds[kWsLdsBytes];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int n0 = blockIdx.y * NCOL;

  // ---- weights of all offsets for columns [n0, n0 + NCOL) -> LDS
  {
    const T* B = reinterpret_cast<const T*>(p.b);
    const int total = p.kv * NCOL * CHUNKS;
    for (int i = tid; i < total; i += blockDim.x) {
      const int c = i % CHUNKS;
      const int n = (i / CHUNKS) % NCOL;
      const int f = i / (CHUNKS * NCOL);
      u32x4 v = zero16();
      if (n0 + n < p.nout) v = ld16(B + (int64_t)f * p.b_fs + (int64_t)(n0 + n) * p.b_ns + c * E);
      const int slot = ((f * NCOL + n) * CHUNKS + (c ^ (n & SWZ))) * 16;
      *reinterpret_cast<u32x4*>(wlds + slot) = v;
    }
  }
  __syncthreads();

  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a, p.a_bytes);
  const uint32_t kvmask = p.kv >= 32 ? 0xffffffffu : ((1u << p.kv) - 1u);
  OutT* Cout = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
  const int gwaves = gridDim.x * kWsWaves;

  for (int t = blockIdx.x * kWsWaves + wave; t < ntiles; t += gwaves) {
    const int row0 = t * kMaskTile;
    const int r = row0 + lr;
    const bool rv = r < p.m;
    uint32_t m = 0;
    int col = 0, orow = 0;
    if (rv) {
      m = p.mask ? p.mask[r] : kvmask;
      col = p.pair_col ? p.pair_col[r] : r;
      orow = p.out_row ? p.out_row[r] : r;
    }
    m |= __shfl_xor(m, 1);
    m |= __shfl_xor(m, 2);
    m |= __shfl_xor(m, 4);
    m |= __shfl_xor(m, 8);
    uint32_t km = __builtin_amdgcn_readfirstlane(m) & p.mask_filter & kvmask;
    if (blockIdx.y == 0 && p.mask_out != nullptr && lane == 0) p.mask_out[t] = km;
    if (p.skip_empty && km == 0u) continue;
    if (p.reverse_mask) km = __builtin_bitreverse32(km) >> (32 - p.kv);

    f32x4 acc[NBC];
#pragma unroll
    for (int nb = 0; nb < NBC; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    uint32_t rem = km;
    while (rem) {
      // this batch's offsets (scalar), missing ones marked -1
      int fo[BATCH];
#pragma unroll
      for (int i = 0; i < BATCH; ++i) {
        fo[i] = rem ? __builtin_ctz(rem) : -1;
        rem &= rem - 1u;
      }
      // lane group lg loads the pair indices of batch slots lg and lg + 4
      int fa = lg == 0 ? fo[0] : lg == 1 ? fo[1] : lg == 2 ? fo[2] : fo[3];
      int fb = -1;
      if constexpr (BATCH == 8) fb = lg == 0 ? fo[4] : lg == 1 ? fo[5] : lg == 2 ? fo[6] : fo[7];
      int ia = -1, ib = -1;
      if (p.pairs) {
        ia = p.pairs[(int64_t)(fa < 0 ? 0 : fa) * p.pair_ld + col];
        if constexpr (BATCH == 8) ib = p.pairs[(int64_t)(fb < 0 ? 0 : fb) * p.pair_ld + col];
      } else {
        ia = col;
        ib = col;
      }
      ia = fa < 0 ? -1 : ia;
      ib = fb < 0 ? -1 : ib;
      // gathered rows of the batch: slot i's index for row lr sits in lane
      // (i % 4) * 16 + lr
      u32x4 af[BATCH][KC];
#pragma unroll
      for (int i = 0; i < BATCH; ++i) {
        const int idx = __shfl(i < 4 ? ia : ib, (i & 3) * 16 + lr);
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
          const int k = kc * CH + lg * E;
          const uint32_t off = idx >= 0 ? (uint32_t)(((int64_t)idx * p.lda + k) * (int64_t)sizeof(T)) : kBufOOB;
          af[i][kc] = bld16(ra, off);
        }
      }
#pragma unroll
      for (int i = 0; i < BATCH; ++i) {
        const int f = fo[i];
        if (f < 0) break;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
          for (int nb = 0; nb < NBC; ++nb) {
            const int n = nb * 16 + lr;
            const int c = kc * (CH / E) + lg;
            const u32x4 bfr =
                *reinterpret_cast<const u32x4*>(wlds + ((f * NCOL + n) * CHUNKS + (c ^ (n & SWZ))) * 16);
            MM::step(acc[nb], af[i][kc], bfr);
          }
        }
      }
    }

    // epilogue: accumulator element i of column block nb is row 4*lg + i,
    // column nb*16 + lr
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = 4 * lg + i;
      const int o = __shfl(orow, rr);
      if (row0 + rr < p.m) {
#pragma unroll
        for (int nb = 0; nb < NBC; ++nb) {
          const int n = n0 + nb * 16 + lr;
          if (n < p.nout) {
            float v = acc[nb][i];
            if (p.scale) v *= p.scale[n];
            v *= p.alpha;
            if (p.beta != 0.f) {
              const float src = p.d_is_bias ? to_f32(D[n]) : to_f32(D[(int64_t)o * p.ldd + n]);
              v += p.beta * src;
            }
            v = apply_act(v, p.act_type, p.act_alpha);
            Cout[(int64_t)o * p.ldc + n] = from_f32<OutT>(v);
          }
        }
      }
    }
  }
}

// Columns per block slice that fit the LDS budget (multiple of 16, <= 64),
// 0 when this kernel does not apply.
template <typename T>
inline int ws_ncol(int kv, int kred) {
  constexpr int CH = Mma<T>::kChunk;
  if (kred % CH != 0 || kred / CH > 4 || kv < 1 || kv > 32) return 0;
  const int64_t per_col = (int64_t)kv * kred * (int64_t)sizeof(T);
  int ncol = (int)(kWsLdsBytes / per_col) / 16 * 16;
  if (ncol > 64) ncol = 64;
  return ncol >= 16 ? ncol : 0;
}

}  // namespace cs

namespace cs {

template <typename T, typename OutT, int KC>
static hipError_t launch_ws_kc(const RowGatherArgs& a, int ncol, hipStream_t s) {
  static int ncu = 0;
  if (ncu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
      ncu = 256;
  }
  const int ntiles = (a.m + kMaskTile - 1) / kMaskTile;
  // one block per CU in total: the column slices share the CUs
  const int slices = (a.nout + ncol - 1) / ncol;
  const int per_slice = std::max(1, std::min(ncu / std::max(slices, 1), (ntiles + kWsWaves - 1) / kWsWaves));
  dim3 grid((unsigned)per_slice, (unsigned)slices, 1);
  dim3 block(64 * kWsWaves, 1, 1);
  if (ntiles == 0 || grid.y == 0) return hipSuccess;
  switch (ncol / 16) {
    case 1: hipLaunchKernelGGL((rowgather_ws_kernel<T, OutT, KC, 1>), grid, block, 0, s, a, ntiles); break;
    case 2: hipLaunchKernelGGL((rowgather_ws_kernel<T, OutT, KC, 2>), grid, block, 0, s, a, ntiles); break;
    case 4: hipLaunchKernelGGL((rowgather_ws_kernel<T, OutT, KC, 4>), grid, block, 0, s, a, ntiles); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Launches the weight-stationary kernel when it applies; *used = false otherwise.
template <typename T, typename OutT>
static hipError_t launch_ws(const RowGatherArgs& a, hipStream_t s, bool* used) {
  *used = false;
  int ncol = ws_ncol<T>(a.kv, a.kred);
  if (ncol == 48) ncol = 32;
  if (ncol == 0 || a.pair_col == nullptr) return hipSuccess;
  const int kc = a.kred / Mma<T>::kChunk;
  *used = true;
  switch (kc) {
    case 1: return launch_ws_kc<T, OutT, 1>(a, ncol, s);
    case 2: return launch_ws_kc<T, OutT, 2>(a, ncol, s);
    case 4: return launch_ws_kc<T, OutT, 4>(a, ncol, s);
    default: *used = false; return hipSuccess;
  }
}

}  // namespace cs
