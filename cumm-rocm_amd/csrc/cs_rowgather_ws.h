// Weight-stationary row-gather kernel (sparse conv forward / dgrad).
//
// Measured on MI355X, a dependent gather round trip under full load costs
// ~1.5-2 us, and a kernel whose waves walk a tile's kernel offsets one round
// trip at a time lasts as long as its heaviest tiles (up to 24 offsets).  This
// kernel bounds every tile to three round trips (masks -> pair indices of a
// batch of offsets -> their gathered rows) and keeps register use at or below
// 128 VGPRs so 16 waves fit on a CU (more waves in flight measured faster than
// deeper per-wave pipelines at 12 waves):
//   * persistent blocks, one per CU: the weights of all kv offsets for a slice
//     of NCOL output columns are loaded once into LDS (<= 108 KiB, 16-byte
//     chunks XOR-swizzled by row so MFMA operand reads are conflict-free);
//   * each wave takes 16-row mask tiles round-robin over the whole grid;
//   * per tile: one load round trip for row masks / argsort, one for the pair
//     indices of a batch of 3 / 3 / 2 offsets (1 / 2 / 4 reduction chunks;
//     each lane group loads a different offset), one for the gathered rows of
//     those offsets (raw buffer
//     loads, index -1 reads zero); MFMAs take weights from LDS.
// Used when kv * NCOL * C * sizeof(T) fits the LDS budget with NCOL >= 16.
#pragma once
#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

constexpr int kWsLdsBytes = 110592;  // 108 KiB of weights per block
constexpr int kWsWaves = 16;         // waves per block (4 per SIMD: <= 128 VGPRs each)
constexpr int kWsStageBytes = 2048;  // per-wave LDS tile for coalesced row stores
constexpr int kWsPoolMax = 4096;     // tiles per block ordered by cost (larger pools keep grid order)

// KC: reduction chunks (C / chunk); NBC: 16-column blocks per wave (NCOL / 16)
template <typename T, typename OutT, int KC, int NBC>
__global__ __launch_bounds__(64 * kWsWaves) void rowgather_ws_kernel(RowGatherArgs p, int ntiles) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int NCOL = NBC * 16;
  constexpr int C = KC * CH;            // reduction length handled (p.kred rounded up to whole chunks)
  constexpr int CHUNKS = C / E;         // 16-byte chunks per weight row
  static_assert((CHUNKS & (CHUNKS - 1)) == 0, "chunk swizzle needs a power-of-two row");
  // offsets per gather round trip (measured on MI355X: 16 waves beat 12 waves
  // with 8-offset batches; at one and two reduction chunks 3 offsets beat 2,
  // 4, 6 and 8, where fewer registers leave the next tile's loads more room
  // to overlap)
  constexpr int BATCH = KC <= 2 ? 3 : 2;
  __shared__ __attribute__((aligned(16))) char wlds[kWsLdsBytes];
  __shared__ __attribute__((aligned(16))) char stage_lds[kWsWaves * kWsStageBytes];
  __shared__ int tile_ctr;  // next entry of this block's tile pool
  __shared__ uint16_t pool[kWsPoolMax];  // pool entries, heaviest tiles first
  __shared__ uint8_t pcost[kWsPoolMax];  // offsets per pool entry
  __shared__ int hist[34];               // pool entries per cost bucket (32 - offsets)

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int n0 = blockIdx.y * NCOL;

  const uint32_t kvmask = p.kv >= 32 ? 0xffffffffu : ((1u << p.kv) - 1u);
  // per-task row data: row mask, pair column (argsort) and output row of the
  // lane's row (lr); rows past m read as empty
  // Row data load as unconditional buffer loads (rows past m read 0), so a
  // prefetch is waited on only where it is used: a load inside a branch, or
  // a select on a loaded value, makes the compiler wait right there for
  // every outstanding load.  The launcher takes this kernel only for masked
  // convs, whose pair column and output row are both the argsort.
  struct Meta { uint32_t m; int col, orow; };
  const uint32_t mbytes = (uint32_t)p.m * 4u;
  const __amdgpu_buffer_rsrc_t rmask = make_rsrc(p.mask, mbytes);
  const __amdgpu_buffer_rsrc_t rcol = make_rsrc(p.pair_col, mbytes);
  auto load_meta = [&](int t, Meta& M) {
    const int r = t * kMaskTile + lr;
    const uint32_t off = t < ntiles && r < p.m ? (uint32_t)r * 4u : kBufOOB;
    M.m = __builtin_amdgcn_raw_buffer_load_b32(rmask, off, 0, 0);
    M.col = (int)__builtin_amdgcn_raw_buffer_load_b32(rcol, off, 0, 0);
    M.orow = M.col;
  };

  // Block b's pool is tiles b, b + G, b + 2G, ... (G = gridDim.x: a uniform
  // sample of the tile order, so pools weigh about the same).  A tile costs
  // about one gather round trip per batch of its offsets, and a wave runs only
  // a few tiles, so the pool is ordered heaviest first (longest-processing-
  // time order) before the waves draw it: entry k's cost is the popcount of
  // its tile mask.  Which wave runs a tile does not change any result.
  const int P = blockIdx.x < ntiles ? (ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
  const bool ordered = P <= kWsPoolMax;
  auto pool_tile = [&](int v) -> int {  // tile of pool position v
    if (v >= P) return ntiles;
    const int k = ordered ? (int)pool[v] : v;
    return (int)blockIdx.x + (int)gridDim.x * k;
  };
  // row data of the first 64 pool entries (16 lanes each), issued before the
  // weights: it orders the pool and is the first tiles' row data
  constexpr int kWsFirst = 64 * kWsWaves / kMaskTile;  // pool entries of the first pass
  static_assert(sizeof(Meta) * 64 * kWsWaves <= kWsWaves * kWsStageBytes, "first-pass row data exceeds the stage buffers");
  Meta M0;
  M0.m = 0u;
  M0.col = 0;
  M0.orow = 0;
  if (ordered) {
    const int k = tid >> 4;  // pool entry k is tile blockIdx.x + G * k (past ntiles when k >= P)
    load_meta(k < P ? (int)blockIdx.x + (int)gridDim.x * k : ntiles, M0);
  }
  int t0 = ntiles, t1 = ntiles;
  Meta Mc, Mn;


  // per-column epilogue constants of this lane's columns, loaded once
  float ep_s[NBC], ep_b[NBC];
#pragma unroll
  for (int nb = 0; nb < NBC; ++nb) {
    const int n = n0 + nb * 16 + lr;
    const bool ok = n < p.nout;
    ep_s[nb] = (p.scale && ok) ? p.scale[n] : 1.f;
    ep_b[nb] = (IsInt8<T>::value && ok) ? p.bias_f[n] : 0.f;
  }

  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a, p.a_bytes);
  OutT* Cout = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
  // coalesced row stores through a per-wave LDS tile (beta == 0, whole slice
  // inside nout, 16-byte aligned rows)
  constexpr int ROWB = NCOL * (int)sizeof(OutT);
  constexpr bool kStage = ROWB % 16 == 0 && 16 * ROWB <= kWsStageBytes;
  const bool vec_store = kStage && p.beta == 0.f && n0 + NCOL <= p.nout &&
                         ((int64_t)p.ldc * (int64_t)sizeof(OutT)) % 16 == 0 &&
                         ((reinterpret_cast<uintptr_t>(Cout) + (uintptr_t)n0 * sizeof(OutT)) & 15u) == 0;
  char* stage = stage_lds + wave * kWsStageBytes;

  // tile mask (also stored to mask_output), filter, reverse
  auto tile_mask = [&](int t, uint32_t m) -> uint32_t {
    m |= __shfl_xor(m, 1);
    m |= __shfl_xor(m, 2);
    m |= __shfl_xor(m, 4);
    m |= __shfl_xor(m, 8);
    uint32_t km = __builtin_amdgcn_readfirstlane(m) & p.mask_filter & kvmask;
    if (blockIdx.y == 0 && p.mask_out != nullptr && lane == 0 && t < ntiles) p.mask_out[t] = km;
    return km;
  };
  auto prepare = [&](uint32_t km, bool& live) -> uint32_t {
    live = !(p.skip_empty && km == 0u);
    if (!live) return 0u;
    return p.reverse_mask ? __builtin_bitreverse32(km) >> (32 - p.kv) : km;
  };
  // i-th set bit of a batch mask (-1 past the end); i is a compile-time slot
  auto slot_f = [](uint32_t bm, int i) -> int {
    for (int j = 0; j < i; ++j) bm &= bm - 1u;
    return bm ? __builtin_ctz(bm) : -1;
  };
  // byte offset of pair[f][col]; f < 0 sets bit 31 (past the extent: reads 0)
  const __amdgpu_buffer_rsrc_t rpair = make_rsrc(p.pairs, (uint32_t)((int64_t)p.kv * p.pair_ld * 4));
  auto pair_off = [&](int f, int col) -> uint32_t {
    return (uint32_t)(((int64_t)f * p.pair_ld + col) * 4) | ((uint32_t)(f >> 31) << 31);
  };
  // next batch: the first BATCH offsets of rem (returned as a bit mask);
  // pair indices issued (lane group lg loads slot lg), not waited on
  auto batch = [&](uint32_t& rem, int col, int& ia) -> uint32_t {
    uint32_t bm = 0u;
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      const uint32_t low = rem & (0u - rem);
      bm |= low;
      rem ^= low;
    }
    int fs[BATCH];
#pragma unroll
    for (int i = 0; i < BATCH; ++i) fs[i] = slot_f(bm, i);
    static_assert(BATCH == 2 || BATCH == 3, "one pair-index load per lane group");
    int fa;
    if constexpr (BATCH == 3) fa = lg == 0 ? fs[0] : lg == 1 ? fs[1] : lg == 2 ? fs[2] : -1;
    else fa = lg == 0 ? fs[0] : lg == 1 ? fs[1] : -1;
    // unconditional buffer load (slot past the batch: offset past the extent)
    ia = (int)__builtin_amdgcn_raw_buffer_load_b32(rpair, pair_off(fa, col), 0, 0);
    return bm;
  };
  // gathered rows of a batch: slot i's index for row lr sits in lane i * 16 + lr
  auto gather = [&](uint32_t bm, int ia, u32x4 (&af)[BATCH][KC]) {
    const int cnt = __builtin_popcount(bm);
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      int idx = __shfl(ia, i * 16 + lr);
      idx = i < cnt ? idx : -1;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        const int k = kc * CH + lg * E;
        // idx < 0 sets bit 31: an offset past the extent, which reads zero;
        // so do reduction elements past kred (a reduction padded to whole
        // chunks).  Arithmetic rather than a select keeps the load out of a
        // branch.
        const uint32_t off = (uint32_t)(((int64_t)idx * p.lda + k) * (int64_t)sizeof(T)) |
                             ((uint32_t)(idx >> 31) << 31) | ((uint32_t)(k >= p.kred) << 31);
        af[i][kc] = bld16(ra, off);
      }
    }
  };
  auto mma_batch = [&](uint32_t bm, const u32x4 (&af)[BATCH][KC], f32x4 (&acc)[NBC]) {
#pragma unroll
    for (int i = 0; i < BATCH; ++i) {
      if (!bm) break;
      const int f = __builtin_ctz(bm);
      bm &= bm - 1u;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
        for (int nb = 0; nb < NBC; ++nb) {
          const int n = nb * 16 + lr;
          const int c = kc * (CH / E) + lg;
          const u32x4 bfr =
              *reinterpret_cast<const u32x4*>(wlds + ((f * NCOL + n) * CHUNKS + (c ^ chunk_swz<CHUNKS>(n))) * 16);
          MM::step(acc[nb], af[i][kc], bfr);
        }
      }
    }
  };
  // epilogue: accumulator element i of column block nb is row 4*lg + i,
  // column nb*16 + lr
  auto epilogue = [&](int t, int orow, const f32x4 (&acc)[NBC]) {
    const int row0 = t * kMaskTile;
    if (vec_store) {
      OutT* st = reinterpret_cast<OutT*>(stage);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int nb = 0; nb < NBC; ++nb) {
          const int n = n0 + nb * 16 + lr;
          float v;
          if constexpr (IsInt8<T>::value) {
            v = i8_epilogue(p, __float_as_int(acc[nb][i]), ep_s[nb], ep_b[nb], n, 0);  // beta == 0 here
          } else {
            v = acc[nb][i];
            if (p.scale) v *= ep_s[nb];
            v = apply_act(v * p.alpha, p.act_type, p.act_alpha);
          }
          st[(4 * lg + i) * NCOL + nb * 16 + lr] = from_f32<OutT>(v);
        }
      __builtin_amdgcn_wave_barrier();
      constexpr int LPR = ROWB / 16;   // lanes per row
      constexpr int RPI = 64 / LPR;    // rows per store instruction (may exceed 16)
#pragma unroll
      for (int rb = 0; rb < 16; rb += RPI) {
        const int rr = rb + lane / LPR;
        const int chunk = lane % LPR;
        const bool ok = rr < 16 && row0 + rr < p.m;
        const int o = __shfl(orow, rr & 15);
        if (ok) {
          const u32x4 v = *reinterpret_cast<const u32x4*>(stage + rr * ROWB + chunk * 16);
          *reinterpret_cast<u32x4*>(reinterpret_cast<char*>(Cout + (int64_t)o * p.ldc + n0) + chunk * 16) = v;
        }
      }
      __builtin_amdgcn_wave_barrier();
      return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = 4 * lg + i;
      const int o = __shfl(orow, rr);
      if (row0 + rr < p.m) {
#pragma unroll
        for (int nb = 0; nb < NBC; ++nb) {
          const int n = n0 + nb * 16 + lr;
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
  };

  // The wave's tiles are software-pipelined: while the gathered rows of tile
  // k are in flight, the pair indices of tile k+1 and the row data of tile
  // k+2 load too, so a tile costs about one round trip.  Tiles with more than
  // BATCH offsets run their further batches with the next batch's pair indices
  // in flight.  The waves take the ordered pool's entries in turn through an
  // LDS counter, so a wave that drew heavy tiles takes fewer.
  auto next_tile = [&]() -> int {
    int v = 0;
    if (lane == 0) v = atomicAdd(&tile_ctr, 1);
    v = __builtin_amdgcn_readfirstlane(v);
    return pool_tile(v);
  };
  // ---- weights of all offsets for columns [n0, n0 + NCOL) -> LDS.  Each
  // thread issues its (at most WPT) loads before writing any of them, so the
  // preload costs one round trip instead of one per chunk.  The first tile's
  // row data loads before the weights and its pair indices are issued while
  // the weights are in flight (loads complete in issue order, so the LDS
  // stores wait for the weights only).
  bool live_c = false;
  uint32_t rem_c = 0u, bm_c = 0u;
  int ia_c = -1;
  auto first_tile = [&]() {
    // order the pool: counting sort by offsets, descending
    if (ordered) {
      if (tid < 34) hist[tid] = 0;
      __syncthreads();
      // the first pass's row data, in the (free until the first epilogue) stage buffers
      Meta* sm = reinterpret_cast<Meta*>(stage_lds);
      sm[tid] = M0;
      for (int k0 = 0; k0 < P; k0 += kWsFirst) {
        const int k = k0 + (tid >> 4);
        uint32_t m = M0.m;
        if (k0 > 0) {
          Meta Mk;
          load_meta(k < P ? (int)blockIdx.x + (int)gridDim.x * k : ntiles, Mk);
          m = Mk.m;
        }
        m |= __shfl_xor(m, 1);
        m |= __shfl_xor(m, 2);
        m |= __shfl_xor(m, 4);
        m |= __shfl_xor(m, 8);
        if ((tid & 15) == 0 && k < P) {
          const int c = __builtin_popcount(m & p.mask_filter & kvmask);
          pcost[k] = (uint8_t)c;
          atomicAdd(&hist[32 - c], 1);
        }
      }
      __syncthreads();
      if (tid < 64) {  // exclusive prefix of the 33 buckets: one wave scan
        const int h = tid < 33 ? hist[tid] : 0;
        int incl = h;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int u = __shfl_up(incl, d);
          if (tid >= d) incl += u;
        }
        if (tid < 33) hist[tid] = incl - h;
      }
      __syncthreads();
      for (int k = tid; k < P; k += 64 * kWsWaves) pool[atomicAdd(&hist[32 - pcost[k]], 1)] = (uint16_t)k;
      __syncthreads();
    }
    // each wave's first two tiles: pool positions wave and kWsWaves + wave;
    // row data of first-pass entries comes from LDS
    t0 = pool_tile(wave);
    t1 = pool_tile(kWsWaves + wave);
    auto first_meta = [&](int v, int t, Meta& M) {
      const int k = ordered && v < P ? (int)pool[v] : kWsFirst;
      if (k < kWsFirst) M = reinterpret_cast<const Meta*>(stage_lds)[k * kMaskTile + lr];
      else load_meta(t, M);
    };
    first_meta(wave, t0, Mc);
    first_meta(kWsWaves + wave, t1, Mn);
    __syncthreads();  // the stage buffers are free again
    rem_c = prepare(tile_mask(t0, Mc.m), live_c);  // empty past the last tile
    bm_c = batch(rem_c, Mc.col, ia_c);
  };
  if (p.b_kmajor) {
    // B element (f, n, k) at b[k * b_ks + f * b_fs + n] (dgrad reads the KRSC
    // weights directly): a unit of E reduction rows x E columns loads as E
    // 16-byte rows and is transposed in registers into E chunks along k.
    // Consecutive threads take consecutive column units of the same rows, so
    // each load instruction reads whole row segments (the column units of a
    // row are contiguous in memory; consecutive reduction units are not).
    constexpr int S = (int)sizeof(T);
    constexpr int NU = NCOL / E, KU = C / E;
    constexpr int UPT = (32 * NU * KU + 64 * kWsWaves - 1) / (64 * kWsWaves);  // kv <= 32
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.b, p.b_bytes);
    const int units = p.kv * NU * KU;
    u32x4 v[UPT][E];
#pragma unroll
    for (int j = 0; j < UPT; ++j) {
      const int u = tid + j * (64 * kWsWaves);
      const int f = u / (NU * KU), ku = (u / NU) % KU, nu = u % NU;
      const bool ok = u < units && n0 + nu * E < p.nout && ku * E < p.kred;  // padded reduction rows read 0
#pragma unroll
      for (int i = 0; i < E; ++i)
        v[j][i] = bld16(rb, ok ? (uint32_t)(((int64_t)(ku * E + i) * p.b_ks + (int64_t)f * p.b_fs + n0 + nu * E) * S)
                             : kBufOOB);
    }
    first_tile();
#pragma unroll
    for (int j = 0; j < UPT; ++j) {
      const int u = tid + j * (64 * kWsWaves);
      if (u < units) {
        const int f = u / (NU * KU), ku = (u / NU) % KU, nu = u % NU;
#pragma unroll
        for (int jn = 0; jn < E; ++jn) {
          u32x4 o = zero16();
#pragma unroll
          for (int i = 0; i < E; ++i) {
            const uint32_t w = v[j][i][(jn * S) >> 2];
            uint32_t e;
            if constexpr (S == 4) {
              e = w;
            } else {
              e = (w >> (((jn * S) & 3) * 8)) & ((1u << (8 * S)) - 1u);
            }
            o[(i * S) >> 2] |= e << (((i * S) & 3) * 8);
          }
          const int n = nu * E + jn;
          *reinterpret_cast<u32x4*>(wlds + ((f * NCOL + n) * CHUNKS + (ku ^ chunk_swz<CHUNKS>(n))) * 16) = o;
        }
      }
    }
  } else {
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.b, p.b_bytes);
    const int total = p.kv * NCOL * CHUNKS;
    constexpr int WPT = (kWsLdsBytes / 16 + 64 * kWsWaves - 1) / (64 * kWsWaves);
    u32x4 v[WPT];
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const int i = tid + j * (64 * kWsWaves);
      const int c = i % CHUNKS;
      const int n = (i / CHUNKS) % NCOL;
      const int f = i / (CHUNKS * NCOL);
      const bool ok = i < total && n0 + n < p.nout && c * E < p.kred;  // padded reduction chunks read 0
      v[j] = bld16(rb, ok ? (uint32_t)(((int64_t)f * p.b_fs + (int64_t)(n0 + n) * p.b_ns + c * E) * (int)sizeof(T))
                          : kBufOOB);
    }
    first_tile();
#pragma unroll
    for (int j = 0; j < WPT; ++j) {
      const int i = tid + j * (64 * kWsWaves);
      if (i < total) {
        const int c = i % CHUNKS;
        const int n = (i / CHUNKS) % NCOL;
        const int f = i / (CHUNKS * NCOL);
        *reinterpret_cast<u32x4*>(wlds + ((f * NCOL + n) * CHUNKS + (c ^ chunk_swz<CHUNKS>(n))) * 16) = v[j];
      }
    }
  }
  if (tid == 0) tile_ctr = 2 * kWsWaves;
  __syncthreads();
  int t = t0;
  if (t >= ntiles) return;
  int tn = t1;
  while (true) {
    u32x4 af[BATCH][KC];
    gather(bm_c, ia_c, af);
    // next tile: its mask is final once its row data has landed
    const int tnn = tn < ntiles ? next_tile() : ntiles;
    bool live_n = false;
    uint32_t rem_n = 0u, bm_n = 0u;
    int ia_n = -1;
    Meta Mnn;
    // unconditional: past the last tile the row data reads empty and the
    // batch is empty, so no load sits behind a branch
    rem_n = prepare(tile_mask(tn, Mn.m), live_n);
    bm_n = batch(rem_n, Mn.col, ia_n);
    load_meta(tnn, Mnn);
    // a tile with more than BATCH offsets: the pair indices of each further
    // batch load while the rows of the batch before it are gathered, so every
    // further batch costs one round trip
    uint32_t bm_x = 0u;
    int ia_x = -1;
    bm_x = batch(rem_c, Mc.col, ia_x);
    f32x4 acc[NBC];
#pragma unroll
    for (int nb = 0; nb < NBC; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    mma_batch(bm_c, af, acc);
    while (bm_x) {
      gather(bm_x, ia_x, af);
      int ia_y = -1;
      const uint32_t bm_y = batch(rem_c, Mc.col, ia_y);
      mma_batch(bm_x, af, acc);
      bm_x = bm_y;
      ia_x = ia_y;
    }
    if (live_c) epilogue(t, Mc.orow, acc);
    if (tn >= ntiles) break;
    t = tn;
    tn = tnn;
    Mc = Mn;
    Mn = Mnn;
    live_c = live_n;
    rem_c = rem_n;
    bm_c = bm_n;
    ia_c = ia_n;
  }
}

// Columns per block slice that fit the LDS budget (multiple of 16, <= 64),
// 0 when this kernel does not apply.
template <typename T>
inline int ws_ncol(int kv, int kred) {
  constexpr int CH = Mma<T>::kChunk;
  constexpr int E = 16 / (int)sizeof(T);
  // a reduction of whole 16-byte rows is padded to whole MFMA chunks (the
  // padding reads zero), e.g. 16 channels of a 16-bit type
  if (kred < 1 || kred % E != 0 || (kred + CH - 1) / CH > 4 || kv < 1 || kv > 32) return 0;
  const int64_t per_col = (int64_t)kv * ((kred + CH - 1) / CH * CH) * (int64_t)sizeof(T);
  int ncol = (int)(kWsLdsBytes / per_col) / 16 * 16;
  if (ncol > 64) ncol = 64;
  return ncol >= 16 ? ncol : 0;
}

}  // namespace cs

namespace cs {

template <typename T, typename OutT, int KC>
static hipError_t launch_ws_kc(const RowGatherArgs& a, int ncol, hipStream_t s) {
  // CUs of the current device, queried per launch (the runtime caches the
  // attribute), so calls on different devices or threads share no state
  int ncu = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
    ncu = 256;
  const int ntiles = (a.m + kMaskTile - 1) / kMaskTile;
  // one block per CU in total: the column slices share the CUs
  const int slices = (a.nout + ncol - 1) / ncol;
  const int per_slice = std::max(1, std::min(ncu / std::max(slices, 1), (ntiles + kWsWaves - 1) / kWsWaves));
  dim3 grid((unsigned)per_slice, (unsigned)slices, 1);
  dim3 block(64 * kWsWaves, 1, 1);
  if (ntiles == 0 || grid.y == 0) return hipSuccess;
  switch (ncol / 16) {
    case 1: CS_LAUNCH((rowgather_ws_kernel<T, OutT, KC, 1>), grid, block, 0, s, a, ntiles); break;
    case 2: CS_LAUNCH((rowgather_ws_kernel<T, OutT, KC, 2>), grid, block, 0, s, a, ntiles); break;
    case 4:
      // four reduction chunks with 64 columns do not fit 128 VGPRs; launch_ws caps them at 32
      if constexpr (KC < 4) {
        CS_LAUNCH((rowgather_ws_kernel<T, OutT, KC, 4>), grid, block, 0, s, a, ntiles);
        break;
      }
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Launches the weight-stationary kernel when it applies; *used = false otherwise.
// The column slice is the largest of 16 / 32 / 64 that fits the LDS budget and
// does not exceed max_ncol.
template <typename T, typename OutT>
static hipError_t launch_ws(const RowGatherArgs& a, int max_ncol, hipStream_t s, bool* used) {
  *used = false;
  int ncol = std::min(ws_ncol<T>(a.kv, a.kred), max_ncol);
  if (ncol == 48) ncol = 32;
  // masked convs only: row data = mask + argsort (pair column = output row)
  if (ncol == 0 || a.pair_col == nullptr || a.mask == nullptr || a.out_row != a.pair_col || a.pairs == nullptr ||
      (int64_t)a.kv * a.pair_ld * 4 >= ((int64_t)1 << 31))
    return hipSuccess;
  const int kc = (a.kred + Mma<T>::kChunk - 1) / Mma<T>::kChunk;
  if (kc == 4 && ncol > 32) ncol = 32;
  *used = true;
  switch (kc) {
    case 1: return launch_ws_kc<T, OutT, 1>(a, ncol, s);
    case 2: return launch_ws_kc<T, OutT, 2>(a, ncol, s);
    case 4: return launch_ws_kc<T, OutT, 4>(a, ncol, s);
    default: *used = false; return hipSuccess;
  }
}

}  // namespace cs
