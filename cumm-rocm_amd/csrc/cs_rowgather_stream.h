// Streamed-weights row-gather kernel (sparse conv forward / dgrad with wide
// layers, e.g. 128 -> 128 channels).
//
// When the weights of all kernel offsets for every output column do not fit
// the LDS, the weight-stationary kernel has to split the columns into slices
// and every slice gathers all input rows again.  Here a block owns 128 sorted
// rows (eight 16-row mask tiles, one per wave) and all output columns, and
// walks the union of its tiles' offset masks in ascending order.  For each
// offset the block stages that offset's weights [nout][kred] in LDS (double
// buffered) and each wave gathers its 16 rows straight into MFMA operand
// registers; a wave whose own tile lacks the offset skips the MFMAs.  The
// weights of the next offset and the rows of the next offset load while the
// current offset's MFMAs run, and the pair indices two offsets ahead.
// A row's sum runs over offsets in ascending order and a missing neighbour
// adds an exact zero, so results are bit-identical to the other row-gather
// kernels.
#pragma once
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

constexpr int kStWaves = 8;                 // waves per block = mask tiles per block
constexpr int kStLdsBytes = 2 * 32768;      // two weight stages of up to 32 KiB

// KC: reduction chunks (kred / chunk); NB: 16-column blocks (nout <= 16 NB);
// MW: mask words per window of offsets (1, or kMaxMaskWords for dynamic
// masks, kv > 32: offsets are walked window by window, each window's words in
// kernel order; a row's sum still runs over its offsets in ascending order)
template <typename T, typename OutT, int KC, int NB, int MW>
__global__ __launch_bounds__(64 * kStWaves) void rowgather_stream_kernel(RowGatherArgs p, int ntiles) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int C = KC * CH;              // reduction length (== p.kred)
  constexpr int NCOL = NB * 16;
  constexpr int CHUNKS = C / E;           // 16-byte chunks per weight row
  static_assert((CHUNKS & (CHUNKS - 1)) == 0, "chunk swizzle needs a power-of-two row");
  constexpr int STAGE = NCOL * CHUNKS * 16;  // bytes of one offset's weights
  static_assert(2 * STAGE <= kStLdsBytes, "weight stage too large");
  constexpr int NT = 64 * kStWaves;
  constexpr int WCH = (NCOL * CHUNKS + NT - 1) / NT;  // 16-byte weight chunks per thread
  __shared__ __attribute__((aligned(16))) char wlds[kStLdsBytes];
  __shared__ uint32_t wmask[kStWaves][1];
  __shared__ uint32_t tmask[MW > 1 ? kStWaves : 1][MW > 1 ? kMaxDynWords : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  // Rows are sorted by mask, and tiles late in that order carry more offsets
  // (the high offsets are the sort's most significant bits), so blocks walk
  // the tile groups from the last one back: the longest blocks start first
  // and the short ones fill in behind them.
  const int grp = (int)gridDim.x - 1 - (int)blockIdx.x;
  const int t = grp * kStWaves + wave;
  const int row0 = t * kMaskTile;
  const int r = row0 + lr;
  const bool rvalid = t < ntiles && r < p.m;

  // ---- this wave's tile: pair column, output row of row lr
  int col = 0, orow = 0;
  if (rvalid) {
    col = p.pair_col ? p.pair_col[r] : r;
    orow = p.out_row ? p.out_row[r] : r;
  }
  uint32_t km[MW];
  uint32_t U[MW];
  bool live;
  if constexpr (MW == 1) {
    uint32_t m = 0u;
    if (rvalid) m = p.mask ? p.mask[r] : 0xffffffffu;
    m |= __shfl_xor(m, 1);
    m |= __shfl_xor(m, 2);
    m |= __shfl_xor(m, 4);
    m |= __shfl_xor(m, 8);
    km[0] = __builtin_amdgcn_readfirstlane(m) & p.mask_filter & kv_word(p.kv, 0);
    if (t < ntiles && p.mask_out != nullptr && lane == 0) p.mask_out[t] = km[0];
    live = t < ntiles && !(p.skip_empty && km[0] == 0u);
    if (!live) km[0] = 0u;
    if (p.reverse_mask && km[0]) reverse_kv(km, p.kv);
    if (lane == 0) wmask[wave][0] = km[0];
    __syncthreads();
    uint32_t u = 0u;
#pragma unroll
    for (int v = 0; v < kStWaves; ++v) u |= wmask[v][0];
    U[0] = __builtin_amdgcn_readfirstlane(u);
  } else {
    // dynamic masks: every OR word of this wave's tile into LDS (lane group
    // lg loads words lg, lg + 4, ...; one round trip), then windows of MW
    // words are taken from there in kernel order
#pragma unroll
    for (int j = 0; j < kMaxDynWords / 4; ++j) {
      const int q = lg + 4 * j;
      uint32_t m = 0u;
      if (rvalid && q < p.mw) m = p.mask ? p.mask[(int64_t)r * p.mw + q] : 0xffffffffu;
      m |= __shfl_xor(m, 1);
      m |= __shfl_xor(m, 2);
      m |= __shfl_xor(m, 4);
      m |= __shfl_xor(m, 8);
      if (lr == 0 && q < p.mw) tmask[wave][q] = m & p.mask_filter & kv_word(p.kv, q);
    }
    __syncthreads();
    uint32_t any = 0u;
    for (int q = 0; q < p.mw; ++q) any |= tmask[wave][q];
    any = __builtin_amdgcn_readfirstlane(any);
    if (t < ntiles && p.mask_out != nullptr)
      for (int q = lane; q < p.mw; q += 64) p.mask_out[(int64_t)t * p.mw + q] = tmask[wave][q];
    // a tile past the end or skipped for being empty has no bits: the union
    // over the block's tiles below needs no live test
    live = t < ntiles && !(p.skip_empty && any == 0u);
  }
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a, p.a_bytes);
  const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.b, p.b_bytes);
  const T* B = reinterpret_cast<const T*>(p.b);

  // weights of offset f -> registers (WCH 16-byte chunks per thread)
  auto load_w = [&](int f, u32x4 (&wv)[WCH]) {
#pragma unroll
    for (int j = 0; j < WCH; ++j) {
      const int i = tid + j * NT;  // chunk id: n * CHUNKS + c
      const int n = i / CHUNKS, c = i % CHUNKS;
      const bool ok = f >= 0 && i < NCOL * CHUNKS && n < p.nout;
      uint32_t off = kBufOOB;
      if (ok) off = (uint32_t)(((int64_t)f * p.b_fs + (int64_t)n * p.b_ns + c * E) * (int64_t)sizeof(T));
      wv[j] = bld16(rb, off);
    }
  };
  auto store_w = [&](int buf, const u32x4 (&wv)[WCH]) {
    char* base = wlds + buf * STAGE;
#pragma unroll
    for (int j = 0; j < WCH; ++j) {
      const int i = tid + j * NT;
      if (i < NCOL * CHUNKS) {
        const int n = i / CHUNKS, c = i % CHUNKS;
        *reinterpret_cast<u32x4*>(base + (n * CHUNKS + (c ^ chunk_swz<CHUNKS>(n))) * 16) = wv[j];
      }
    }
  };
  // pair index of row lr for window-local offset f (every lane group holds its
  // row's copy); none is read for an offset this wave's tile lacks, whose
  // MFMAs it skips
  auto load_idx = [&](int base, int f) -> int {
    if (f < 0 || !rvalid || !mask_bit(km, f)) return -1;
    return p.pairs ? p.pairs[(int64_t)(base + f) * p.pair_ld + col] : col;
  };
  // gathered operand: lane (lr, lg) holds row lr, reduction elements
  // kc*CH + lg*E .. + E of each chunk kc
  auto gather = [&](int idx, u32x4 (&af)[KC]) {
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const int k = kc * CH + lg * E;
      const uint32_t off = idx >= 0 ? (uint32_t)(((int64_t)idx * p.lda + k) * (int64_t)sizeof(T)) : kBufOOB;
      af[kc] = bld16(ra, off);
    }
  };

  f32x4 acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  // ---- pipeline over the offsets base + f of a window, f0 < f1 < f2 < ...
  // (window-local f from U; the wave's MFMAs run where its km has f)
  auto run = [&](int base) {
  const int f0 = next_offset(U, -1);
  if (f0 >= 0) {
    const int f1 = next_offset(U, f0);
    u32x4 wv[WCH];
    load_w(base + f0, wv);
    int i0 = load_idx(base, f0);
    int i1 = load_idx(base, f1);
    u32x4 a_cur[KC];
    gather(i0, a_cur);
    store_w(0, wv);
    load_w(f1 >= 0 ? base + f1 : -1, wv);  // weights of f1, stored during f0's turn
    __syncthreads();
    int fc = f0, fn = f1, buf = 0;
    while (fc >= 0) {
      const int fnn = fn >= 0 ? next_offset(U, fn) : -1;
      // rows of the next offset and pair indices of the one after
      u32x4 a_nxt[KC];
      gather(i1, a_nxt);
      const int i2 = load_idx(base, fnn);
      if (mask_bit(km, fc)) {
        const char* wb = wlds + buf * STAGE;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) {
            const int n = nb * 16 + lr;
            const int c = kc * (CH / E) + lg;
            const u32x4 bfr = *reinterpret_cast<const u32x4*>(wb + (n * CHUNKS + (c ^ chunk_swz<CHUNKS>(n))) * 16);
            MM::step(acc[nb], a_cur[kc], bfr);
          }
        }
      }
      if (fn >= 0) store_w(buf ^ 1, wv);
      if (fnn >= 0) load_w(base + fnn, wv);
      __syncthreads();  // next stage's weights are in LDS; this stage's reads are done
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) a_cur[kc] = a_nxt[kc];
      i1 = i2;
      fc = fn;
      fn = fnn;
      buf ^= 1;
    }
  }
  };
  if constexpr (MW == 1) {
    run(0);
  } else {
    const bool rev = p.reverse_mask != 0;
    auto tile_word = [&](int q) -> uint32_t { return q >= 0 && q < p.mw ? tmask[wave][q] : 0u; };
    auto union_word = [&](int q) -> uint32_t {
      uint32_t u = 0u;
      if (q >= 0 && q < p.mw) {
#pragma unroll
        for (int v = 0; v < kStWaves; ++v) u |= tmask[v][q];
      }
      return u;
    };
    for (int win = 0; 32 * MW * win < p.kv; ++win) {
#pragma unroll
      for (int i = 0; i < MW; ++i) {
        km[i] = live ? __builtin_amdgcn_readfirstlane(kernel_word(tile_word, MW * win + i, p.kv, rev)) : 0u;
        U[i] = __builtin_amdgcn_readfirstlane(kernel_word(union_word, MW * win + i, p.kv, rev));
      }
      run(32 * MW * win);
    }
  }

  // ---- epilogue: accumulator element i of column block nb is row 4*lg + i,
  // column nb*16 + lr
  if (!live) return;
  // per-column constants of this lane's columns, loaded together after the
  // main loop (one round trip, no registers held across it)
  float ep_s[NB], ep_b[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int n = nb * 16 + lr;
    const bool ok = n < p.nout;
    ep_s[nb] = (p.scale && ok) ? p.scale[n] : 1.f;
    ep_b[nb] = (IsInt8<T>::value && ok) ? p.bias_f[n] : 0.f;
  }
  OutT* Cout = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
  // Whole 16-byte row segments through a per-wave LDS tile (the weight
  // stages are free once the offset walk is done): beta == 0, all columns
  // inside nout, 16-byte aligned rows.
  constexpr int ROWB = NCOL * (int)sizeof(OutT);
  constexpr bool kStage = ROWB % 16 == 0 && kStWaves * 16 * ROWB <= kStLdsBytes;
  const bool vec_store = kStage && p.beta == 0.f && NCOL <= p.nout &&
                         ((int64_t)p.ldc * (int64_t)sizeof(OutT)) % 16 == 0 &&
                         (reinterpret_cast<uintptr_t>(Cout) & 15u) == 0;
  if (vec_store) {
    char* stage = wlds + wave * 16 * ROWB;
    OutT* st = reinterpret_cast<OutT*>(stage);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int n = nb * 16 + lr;
        float v;
        if constexpr (IsInt8<T>::value) {
          v = i8_epilogue(p, __float_as_int(acc[nb][i]), ep_s[nb], ep_b[nb], n, 0);  // beta == 0 here
        } else {
          v = acc[nb][i];
          if (p.scale) v *= ep_s[nb];
          v = apply_act(v * p.alpha, p.act_type, p.act_alpha);
        }
        st[(4 * lg + i) * NCOL + n] = from_f32<OutT>(v);
      }
    __builtin_amdgcn_wave_barrier();
    constexpr int LPR = ROWB / 16;  // lanes per row
    constexpr int RPI = 64 / LPR;   // rows per store instruction
#pragma unroll
    for (int rb = 0; rb < 16; rb += RPI) {
      const int rr = rb + lane / LPR;
      const int chunk = lane % LPR;
      const int o = __shfl(orow, rr & 15);
      if (rr < 16 && row0 + rr < p.m) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(stage + rr * ROWB + chunk * 16);
        *reinterpret_cast<u32x4*>(reinterpret_cast<char*>(Cout + (int64_t)o * p.ldc) + chunk * 16) = v;
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rr = 4 * lg + i;
    const int o = __shfl(orow, rr);
    if (row0 + rr < p.m) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int n = nb * 16 + lr;
        if (n < p.nout) {
          float v;
          if constexpr (IsInt8<T>::value) {
            v = i8_epilogue(p, __float_as_int(acc[nb][i]), ep_s[nb], ep_b[nb], n, o);
          } else {
            v = acc[nb][i];
            if (p.scale) v *= ep_s[nb];
            v *= p.alpha;
            if (p.beta != 0.f) {
              const float src = p.d_is_bias ? to_f32(D[n]) : to_f32(D[(int64_t)o * p.ldd + n]);
              v += p.beta * src;
            }
            v = apply_act(v, p.act_type, p.act_alpha);
          }
          Cout[(int64_t)o * p.ldc + n] = from_f32<OutT>(v);
        }
      }
    }
  }
}

// Launches the streamed-weights kernel when it applies (n-major weights, the
// reduction a multiple of the MFMA chunk with at most 4 chunks, at most 128
// output columns); *used = false otherwise.
template <typename T, typename OutT, int MW>
static hipError_t launch_stream(const RowGatherArgs& a, hipStream_t s, bool* used) {
  *used = false;
  constexpr int CH = Mma<T>::kChunk;
  if (a.b_kmajor || a.pair_col == nullptr || a.kred % CH != 0 || a.kred / CH > 4 || a.nout > 128 || a.nout < 1)
    return hipSuccess;
  const int kc = a.kred / CH;
  const int nb = a.nout <= 32 ? 2 : (a.nout <= 64 ? 4 : 8);
  if (2 * nb * 16 * a.kred * (int)sizeof(T) > kStLdsBytes) return hipSuccess;
  const int ntiles = (a.m + kMaskTile - 1) / kMaskTile;
  if (ntiles == 0) {
    *used = true;
    return hipSuccess;
  }
  dim3 grid((unsigned)((ntiles + kStWaves - 1) / kStWaves)), block(64 * kStWaves);
  *used = true;
#define CS_ST(KC_, NB_) CS_LAUNCH((rowgather_stream_kernel<T, OutT, KC_, NB_, MW>), grid, block, 0, s, a, ntiles)
  switch (kc * 16 + nb) {
    case 1 * 16 + 2: CS_ST(1, 2); break;
    case 1 * 16 + 4: CS_ST(1, 4); break;
    case 1 * 16 + 8: CS_ST(1, 8); break;
    case 2 * 16 + 2: CS_ST(2, 2); break;
    case 2 * 16 + 4: CS_ST(2, 4); break;
    case 2 * 16 + 8: CS_ST(2, 8); break;
    case 4 * 16 + 2: CS_ST(4, 2); break;
    case 4 * 16 + 4: CS_ST(4, 4); break;
    case 4 * 16 + 8: CS_ST(4, 8); break;
    default: *used = false; return hipSuccess;
  }
#undef CS_ST
  return hipGetLastError();
}

}  // namespace cs
