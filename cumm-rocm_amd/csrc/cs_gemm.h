// Gather GEMM kernel of matmul2 (NoShuffle and ShuffleAC: gather A rows,
// scatter C rows).
//
// A block of kGmWaves waves keeps a slab of B (NCOL output columns x KS
// reduction elements, NCOL = 16 * NB, KS = KC MFMA chunks) in LDS and walks
// groups of 16 * kGmWaves GEMM rows, one 16-row tile per wave.  A wave loads
// its rows' A fragments straight from HBM into MFMA operand registers (16
// bytes per lane, raw buffer loads: rows past m and reduction elements past
// k read zeros), multiplies them with B fragments read from LDS, and writes
// its 16 output rows through a per-wave LDS tile as whole-row 16-byte stores
// (rows are scattered by c_rows, so a row is the unit of coalescing).  B is
// the MFMA row operand, so the accumulators hold C transposed and each lane
// owns four consecutive columns of one row, written to LDS as one packed
// store.
//
// When the reduction fits one slab (k <= 8 chunks: 128 f32, 256 f16/bf16;
// 4 chunks, 256, for int8), B is staged once per block; the first tile's A loads are issued
// before B is written to LDS, and every later tile's A loads are in flight
// while the previous tile's MFMAs run.  Longer reductions re-stage B per slab
// and group.  Each output element sums its reduction chunks in ascending
// order with FP32 (int32 for int8) accumulation.
// B is read n-major (b[n*ldb + k]) or k-major (b[k*ldb + n], a row-major
// [k, n] matrix, transposed in registers on its way into LDS), so no
// transposed copy of B is made.
#pragma once
#include <algorithm>
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

constexpr int kGmWaves = 8;
constexpr int kGmRows = 16 * kGmWaves;  // GEMM rows per block step

template <typename T, int NB, int KC>
struct GmGeom {
  static constexpr int S = (int)sizeof(T);
  static constexpr int E = Mma<T>::kElems;   // elements per 16-byte unit
  static constexpr int CH = Mma<T>::kChunk;  // reduction elements per chunk
  static constexpr int NCOL = 16 * NB;
  static constexpr int KS = KC * CH;         // reduction elements per slab
  static constexpr int UNITS = KS / E;       // 16-byte units per B row of a slab
  static constexpr int BLDS = NCOL * UNITS * 16;
  // B staging per thread: 16-byte units along k (n-major b) or along n
  // (k-major b); both layouts hold NCOL * UNITS units per slab
  static constexpr int BN_PER = (NCOL * UNITS + 64 * kGmWaves - 1) / (64 * kGmWaves);
};

// KMAJ: b is k-major (p.b_kmajor); MULTI: the reduction takes several slabs.
// Both are template parameters so that each kernel holds only the registers
// of its own staging and loop.  Single-slab kernels of up to 4 chunks are held
// to 128 VGPRs, so two blocks (16 waves) share a CU: one tile per wave is
// then in flight for most of the work, and more waves hide more latency.
template <typename T, typename OutT, int NB, int KC, bool KMAJ, bool MULTI>
__global__ __launch_bounds__(64 * kGmWaves) __attribute__((amdgpu_waves_per_eu(KC <= 4 && !MULTI ? 4 : 2))) void gemm_rows_kernel(GemmRowsArgs p) {
  using MM = Mma<T>;
  using G = GmGeom<T, NB, KC>;
  constexpr int S = G::S, E = G::E, CH = G::CH, NCOL = G::NCOL, KS = G::KS, UNITS = G::UNITS;
  constexpr int ROWB = NCOL * (int)sizeof(OutT);
  constexpr int OPITCH = ROWB + 32;  // padded row pitch of the output tile
  __shared__ __attribute__((aligned(16))) char blds[G::BLDS];
  __shared__ __attribute__((aligned(16))) char olds[kGmWaves][16 * OPITCH];
  __shared__ int ocl[kGmWaves][2][16];
  __shared__ OutT bias_l[NCOL];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int n0 = blockIdx.y * NCOL;
  const int ngroups = (p.m + kGmRows - 1) / kGmRows;
  if ((int)blockIdx.x >= ngroups) return;  // whole block leaves together
  const int gstep = gridDim.x;

  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a, p.a_bytes);
  const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.b, p.b_bytes);

  // ---- B slab staging: global loads into registers, then LDS (row n, unit u
  // at slot u ^ swizzle(n), so the 16 rows an MFMA operand read touches are
  // spread over the banks)
  u32x4 bst[G::BN_PER];
  auto b_load = [&](int slab) {
#pragma unroll
    for (int j = 0; j < G::BN_PER; ++j) {
      const int i = tid + j * 64 * kGmWaves;
      bool ok = i < NCOL * UNITS;
      uint32_t off;
      if constexpr (!KMAJ) {
        // unit u of row n: reduction elements slab*KS + u*E .. + E
        const int n = i / UNITS, u = i % UNITS;
        const int kk = slab * KS + u * E;
        ok = ok && n0 + n < p.nn && kk < p.k;
        off = (uint32_t)(((int64_t)(n0 + n) * p.ldb + kk) * S);
      } else {
        // unit nu of reduction row kl: columns n0 + nu*E .. + E.  Consecutive
        // threads take consecutive rows of one column unit, so their element
        // stores below fall into one 128-byte span of an LDS row (a column-unit
        // order would put 16 threads on one bank)
        const int kl = i % KS, nu = i / KS;
        const int kk = slab * KS + kl;
        ok = ok && n0 + nu * E < p.nn && kk < p.k;
        off = (uint32_t)(((int64_t)kk * p.ldb + n0 + nu * E) * S);
      }
      bst[j] = bld16(rb, ok ? off : kBufOOB);
    }
  };
  auto b_store = [&]() {
#pragma unroll
    for (int j = 0; j < G::BN_PER; ++j) {
      const int i = tid + j * 64 * kGmWaves;
      if (i < NCOL * UNITS) {
        if constexpr (!KMAJ) {
          const int n = i / UNITS, u = i % UNITS;
          *reinterpret_cast<u32x4*>(blds + (n * UNITS + (u ^ chunk_swz<UNITS>(n))) * 16) = bst[j];
        } else {
          // element e of the unit is B(nu*E + e, kl): written into row nu*E + e
          // of the n-major tile, unit kl / E, position kl % E
          const int kl = i % KS, nu = i / KS;
          const int u = kl / E, pos = (kl % E) * S;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int n = nu * E + e;
            const uint32_t w = bst[j][(e * S) >> 2];
            char* dst = blds + (n * UNITS + (u ^ chunk_swz<UNITS>(n))) * 16 + pos;
            if constexpr (S == 4) {
              *reinterpret_cast<uint32_t*>(dst) = w;
            } else if constexpr (S == 2) {
              *reinterpret_cast<uint16_t*>(dst) = (uint16_t)(w >> ((e & 1) * 16));
            } else {
              *reinterpret_cast<uint8_t*>(dst) = (uint8_t)(w >> ((e & 3) * 8));
            }
          }
        }
      }
    }
  };

  // ---- per-tile row data: A row (gather) and C row (scatter) of this lane's
  // tile row lr; rows past m get -1 (their loads read zeros, nothing stored)
  auto load_rows = [&](int g, int& ia, int& oc) {
    const int r = g * kGmRows + wave * 16 + lr;
    const bool ok = g < ngroups && r < p.m;
    ia = ok ? (p.a_rows ? p.a_rows[r] : r) : -1;
    oc = ok ? (p.c_rows ? p.c_rows[r] : r) : -1;
  };
  auto load_a = [&](int ia, int slab, u32x4 (&af)[KC]) {
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const int kk = slab * KS + kc * CH + lg * E;
      const bool ok = ia >= 0 && kk < p.k;
      af[kc] = bld16(ra, ok ? (uint32_t)(((int64_t)ia * p.lda + kk) * S) : kBufOOB);
    }
  };
  auto mma = [&](const u32x4 (&af)[KC], f32x4 (&acc)[NB]) {
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int n = nb * 16 + lr;
        const u32x4 bfr =
            *reinterpret_cast<const u32x4*>(blds + (n * UNITS + ((kc * 4 + lg) ^ chunk_swz<UNITS>(n))) * 16);
        MM::step(acc[nb], bfr, af[kc]);  // B as the row operand: acc holds C^T tiles
      }
  };

  OutT* Cout = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
  // epilogue source: none, a bias (per column, staged in LDS once per block)
  // or rows of D (staged into the wave's output tile before the values are
  // formed), so every source value is an LDS read and no global load sits in
  // the value computation
  const bool has_src = IsInt8<T>::value ? p.ibeta != 0 : p.beta != 0.f;
  const bool src_rows = has_src && !p.d_is_bias;
  if (has_src && p.d_is_bias) {
    for (int c = tid; c < NCOL; c += 64 * kGmWaves)
      bias_l[c] = n0 + c < p.nn ? D[n0 + c] : OutT{};
  }
  // output element of accumulator value x (float, or int32 bits for int8)
  // with source value sv (0 when there is none)
  auto value = [&](float x, OutT sv) -> OutT {
    if constexpr (IsInt8<T>::value) {
      return int_out<OutT>(int_epilogue(p, __float_as_int(x), to_i32(sv)));
    } else {
      float v = x * p.alpha;
      if (p.beta != 0.f) v += p.beta * to_f32(sv);
      return from_f32<OutT>(apply_act(v, p.act_type, p.act_alpha));
    }
  };
  // With B as the MFMA row operand, lane (lr, lg) holds row lr of the tile
  // and, for column block nb, the four consecutive columns nb*16 + 4*lg + r.
  // Rows leave through the wave's LDS tile: each lane writes its four values
  // as one packed store (row pitch padded by 32 bytes, so a store instruction
  // touches every bank twice at most), then whole 16-byte row pieces go out,
  // 16 lanes per 256 bytes of a row; slices past nn or unaligned rows go out
  // element by element.  Rows of D used as the source come in the same way.
  const bool vec_c = ROWB % 16 == 0 && n0 + NCOL <= p.nn && ((p.ldc * (int64_t)sizeof(OutT)) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(p.c) & 15) == 0;
  const bool vec_d = ROWB % 16 == 0 && n0 + NCOL <= p.nn && ((p.ldd * (int64_t)sizeof(OutT)) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(p.d) & 15) == 0;
  constexpr int LPR = ROWB / 16;  // lanes per row of 16-byte pieces
  constexpr int RPI = 64 / LPR;   // rows per instruction
  // orow: the tile's 16 C rows (-1: no row)
  auto epilogue = [&](const int* orow, const f32x4 (&acc)[NB]) {
    char* st = olds[wave];
    if (src_rows) {
      if (vec_d) {
#pragma unroll
        for (int rb = 0; rb < 16; rb += RPI) {
          const int rr = rb + lane / LPR;
          const int o = rr < 16 ? orow[rr] : -1;
          if (o >= 0) {
            const int chunk = lane % LPR;
            *reinterpret_cast<u32x4*>(st + rr * OPITCH + chunk * 16) =
                *reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(D + (int64_t)o * p.ldd + n0) + chunk * 16);
          }
        }
      } else {
#pragma unroll 1
        for (int rr = 0; rr < 16; ++rr) {
          const int o = orow[rr];
          if (o < 0) continue;
#pragma unroll 1
          for (int c = lane; c < NCOL; c += 64)
            if (n0 + c < p.nn)
              *reinterpret_cast<OutT*>(st + rr * OPITCH + c * (int)sizeof(OutT)) = D[(int64_t)o * p.ldd + n0 + c];
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int cl = nb * 16 + 4 * lg;  // first of this lane's four columns in the slice
      char* slot = st + lr * OPITCH + cl * (int)sizeof(OutT);
      OutT sv[4] = {OutT{}, OutT{}, OutT{}, OutT{}};
      if (src_rows) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sv[r] = reinterpret_cast<const OutT*>(slot)[r];
      } else if (has_src) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sv[r] = bias_l[cl + r];
      }
      OutT v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = value(acc[nb][r], sv[r]);
      if constexpr (sizeof(OutT) == 4) {
        *reinterpret_cast<u32x4*>(slot) = *reinterpret_cast<const u32x4*>(v);
      } else if constexpr (sizeof(OutT) == 2) {
        *reinterpret_cast<uint2*>(slot) = *reinterpret_cast<const uint2*>(v);
      } else {
        *reinterpret_cast<uint32_t*>(slot) = *reinterpret_cast<const uint32_t*>(v);
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (vec_c) {
#pragma unroll
      for (int rb = 0; rb < 16; rb += RPI) {
        const int rr = rb + lane / LPR;
        const int o = rr < 16 ? orow[rr] : -1;
        if (o >= 0) {
          const int chunk = lane % LPR;
          const u32x4 v = *reinterpret_cast<const u32x4*>(st + rr * OPITCH + chunk * 16);
          *reinterpret_cast<u32x4*>(reinterpret_cast<char*>(Cout + (int64_t)o * p.ldc + n0) + chunk * 16) = v;
        }
      }
    } else {
#pragma unroll 1
      for (int rr = 0; rr < 16; ++rr) {
        const int o = orow[rr];
        if (o < 0) continue;
#pragma unroll 1
        for (int c = lane; c < NCOL; c += 64)
          if (n0 + c < p.nn)
            Cout[(int64_t)o * p.ldc + n0 + c] = *reinterpret_cast<const OutT*>(st + rr * OPITCH + c * (int)sizeof(OutT));
      }
    }
    __builtin_amdgcn_wave_barrier();
  };

  if constexpr (!MULTI) {
    // ---- one slab: B staged once.  Step t computes tile t (A in a0 / a1,
    // alternating) while tile t+1's A loads are in flight; the row data of
    // tile t+2 loads at the same time.  The C rows of tiles t and t+1 sit in
    // the wave's two ocl slots (LDS), so no register holding a load in flight
    // is ever copied.
    int ia, oc;
    int g = blockIdx.x;
    load_rows(g, ia, oc);
    b_load(0);
    u32x4 a0[KC], a1[KC];
    load_a(ia, 0, a0);
    ocl[wave][0][lr] = oc;
    load_rows(g + gstep, ia, oc);
    b_store();
    __syncthreads();
    int cur = 0;
    while (true) {
      f32x4 acc[NB];
      load_a(ia, 0, a1);
      ocl[wave][cur ^ 1][lr] = oc;
      load_rows(g + 2 * gstep, ia, oc);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
      mma(a0, acc);
      epilogue(ocl[wave][cur], acc);
      cur ^= 1;
      g += gstep;
      if (g >= ngroups) break;
      load_a(ia, 0, a0);
      ocl[wave][cur ^ 1][lr] = oc;
      load_rows(g + 2 * gstep, ia, oc);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
      mma(a1, acc);
      epilogue(ocl[wave][cur], acc);
      cur ^= 1;
      g += gstep;
      if (g >= ngroups) break;
    }
  } else {
    // ---- several slabs: B re-staged per slab and group (every wave of the
    // block walks the same groups, so the barriers match)
    for (int g = blockIdx.x; g < ngroups; g += gstep) {
      int ia, oc;
      load_rows(g, ia, oc);
      ocl[wave][0][lr] = oc;
      f32x4 acc[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
      const int nslab = (p.k + KS - 1) / KS;
      for (int sl = 0; sl < nslab; ++sl) {
        u32x4 af[KC];
        b_load(sl);
        load_a(ia, sl, af);
        __syncthreads();  // the previous slab's MFMAs are done with the tile
        b_store();
        __syncthreads();
        mma(af, acc);
      }
      epilogue(ocl[wave][0], acc);
    }
  }
}

// Launch shape: NB from the column count (one slice of up to 128 columns per
// grid row), KC the smallest power of two of chunks covering k (8 at most;
// longer reductions take several slabs); blocks per slice from the CUs and
// the blocks per CU that the LDS allows.
template <typename T, typename OutT, int NB, int KC, bool KMAJ, bool MULTI>
static hipError_t launch_gm_kernel(const GemmRowsArgs& a, hipStream_t s) {
  using G = GmGeom<T, NB, KC>;
  int ncu = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    ncu = 256;
  // resident blocks per CU (registers and LDS), queried once per kernel
  static const int per_cu = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, gemm_rows_kernel<T, OutT, NB, KC, KMAJ, MULTI>, 64 * kGmWaves, 0) !=
            hipSuccess || b < 1)
      b = 1;
    return b;
  }();
  const int slices = (a.nn + G::NCOL - 1) / G::NCOL;
  const int ngroups = (a.m + kGmRows - 1) / kGmRows;
  const int per_slice = std::max(1, std::min(ngroups, per_cu * ncu / std::max(slices, 1)));
  if (ngroups == 0 || slices == 0) return hipSuccess;
  CS_LAUNCH((gemm_rows_kernel<T, OutT, NB, KC, KMAJ, MULTI>), dim3((unsigned)per_slice, (unsigned)slices, 1),
                     dim3(64 * kGmWaves, 1, 1), 0, s, a);
  return hipGetLastError();
}

template <typename T, typename OutT, int NB, int KC>
static hipError_t launch_gm_nk(const GemmRowsArgs& a, hipStream_t s, bool multi) {
  if (a.b_kmajor) {
    return multi ? launch_gm_kernel<T, OutT, NB, KC, true, true>(a, s)
                 : launch_gm_kernel<T, OutT, NB, KC, true, false>(a, s);
  }
  return multi ? launch_gm_kernel<T, OutT, NB, KC, false, true>(a, s)
               : launch_gm_kernel<T, OutT, NB, KC, false, false>(a, s);
}

template <typename T, typename OutT, int NB>
static hipError_t launch_gm_n(const GemmRowsArgs& a, hipStream_t s) {
  constexpr int CH = Mma<T>::kChunk;
  // int8 slabs stop at 4 chunks (256 elements): 8 would not fit the registers
  constexpr int KMAX = IsInt8<T>::value ? 4 : 8;
  const int chunks = (a.k + CH - 1) / CH;
  if (chunks <= 1) return launch_gm_nk<T, OutT, NB, 1>(a, s, false);
  if (chunks <= 2) return launch_gm_nk<T, OutT, NB, 2>(a, s, false);
  if (chunks <= 4) return launch_gm_nk<T, OutT, NB, 4>(a, s, false);
  if (KMAX == 8 && chunks <= 8) return launch_gm_nk<T, OutT, NB, KMAX>(a, s, false);
  return launch_gm_nk<T, OutT, NB, KMAX>(a, s, true);
}

template <typename T, typename OutT>
static hipError_t launch_gm(const GemmRowsArgs& a, hipStream_t s) {
  if (a.m < 0 || a.nn < 0 || a.k < 0) return hipErrorInvalidValue;
  if (a.m == 0 || a.nn == 0) return hipSuccess;
  if (a.nn <= 16) return launch_gm_n<T, OutT, 1>(a, s);
  if (a.nn <= 32) return launch_gm_n<T, OutT, 2>(a, s);
  if (a.nn <= 64) return launch_gm_n<T, OutT, 4>(a, s);
  return launch_gm_n<T, OutT, 8>(a, s);
}

}  // namespace cs
