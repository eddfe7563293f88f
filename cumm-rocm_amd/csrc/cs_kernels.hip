// gfx950 kernels for the gather-GEMM / mask-sparse implicit-GEMM path.
//
// rowgather_kernel  : (cs_rowgather.h, instantiated per input dtype in
//                     rg_<dtype>.hip) forward / dgrad, ShuffleAC / dense GEMM.
// wgrad_*          : sparse-conv wgrad / ShuffleAB GEMM.  Per offset, a pair
//                     list (a row, b row) in original row order is built from
//                     the pair table and the effective row masks; the main
//                     kernel stages gathered rows of both operands pair-major
//                     in LDS, reads MFMA operands along the pair dimension
//                     with LDS transpose reads, and writes an FP32 partial per
//                     piece of the list.
// splitk_reduce     : fixed pairwise tree over the split partials + epilogue.
#include <climits>
#include <cstdint>

#include "cs_device.h"
#include "cs_kernels.h"

namespace cs {

// ============================================================ wgrad
// Stage tiles are pair-major in LDS: row = pair, the tile's channels
// contiguous (pitch P bytes), filled by coalesced 16-byte row loads.  Mma<T>
// wants each lane's operand elements along the pair (reduction) dimension:
// lane (i = lane & 15, g = lane >> 4) of the fragment at channel c0 holds E
// elements T[k(g, e)][c0 + i], where k(g, e) is any fixed one-to-one map of
// (lane group, element) onto the chunk's pairs (A and B use the same map).
// gfx950's LDS transpose reads deliver this from pair-major rows: within each
// 16-lane group, ds_read_b64_tr_b16 gives lane i element e = the (i % 4)-th
// value read by lane (i / 4) + 4e, and ds_read_b64_tr_b8 gives lane i byte e =
// the (i % 8)-th byte read by lane (i / 8) + 2e; so lane j points at pair row
// kb + j/4 (j/2 for bytes), channels c0 + 4*(j % 4) (c0 + 8*(j % 2)).
//
// LDS has 64 banks of 4 bytes.  16-bit and 8-bit tiles are unpadded and their
// 16-byte chunks sit at slot chunk ^ swz(row): with the row maps below the 32
// lanes that the LDS serves together (two lane groups) read rows kb .. kb + 7
// (16-bit) or kb .. kb + 15 (8-bit) of one chunk pair or chunk, and swz puts
// them on distinct banks; row stores (16 lanes together) stay conflict-free.
// FP32 tiles keep 16 bytes of row padding (their 16x16x4 reads are plain
// 4-byte loads from four rows 4 apart, conflict-free at that pitch).
typedef short wg_v4i16 __attribute__((ext_vector_type(4)));
typedef int wg_v2i32 __attribute__((ext_vector_type(2)));
#define WG_LDS __attribute__((address_space(3)))

template <typename T> struct WgFrag;
template <> struct WgFrag<float> {
  static constexpr int kPad = 16;
  template <int CPR> __device__ __forceinline__ static int swz(int) { return 0; }
  // 16x16x4 f32: element e of lane (i, g) feeds MFMA e at reduction index 4g + e
  template <int CPR>
  __device__ __forceinline__ static u32x4 load(const char* lds, int pitch, int k0, int c0, int lane) {
    const char* q = lds + (k0 + 4 * (lane >> 4)) * pitch + (c0 + (lane & 15)) * 4;
    return u32x4{*reinterpret_cast<const uint32_t*>(q), *reinterpret_cast<const uint32_t*>(q + pitch),
                 *reinterpret_cast<const uint32_t*>(q + 2 * pitch), *reinterpret_cast<const uint32_t*>(q + 3 * pitch)};
  }
};
template <> struct WgFrag<_Float16> {
  static constexpr int kPad = 0;
  // CPR = 8 or 16 chunks per row
  template <int CPR> __device__ __forceinline__ static int swz(int r) { return (r & 7) * (CPR / 8); }
  // elements 0-3 of lane group g: pairs k0 + 4g + e; elements 4-7: k0 + 16 + 4g + e
  template <int CPR>
  __device__ __forceinline__ static u32x4 load(const char* lds, int pitch, int k0, int c0, int lane) {
    const int j = lane & 15, g = lane >> 4;
    const int ch = c0 + 4 * (j & 3);
    const int r0 = k0 + 4 * g + (j >> 2), r1 = r0 + 16;
    const char* q0 = lds + r0 * pitch + (((ch >> 3) ^ swz<CPR>(r0)) << 4) + (ch & 7) * 2;
    const char* q1 = lds + r1 * pitch + (((ch >> 3) ^ swz<CPR>(r1)) << 4) + (ch & 7) * 2;
    const wg_v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((WG_LDS wg_v4i16*)(q0));
    const wg_v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((WG_LDS wg_v4i16*)(q1));
    const wg_v2i32 l2 = __builtin_bit_cast(wg_v2i32, lo), h2 = __builtin_bit_cast(wg_v2i32, hi);
    return u32x4{(uint32_t)l2.x, (uint32_t)l2.y, (uint32_t)h2.x, (uint32_t)h2.y};
  }
};
template <> struct WgFrag<bf16> : WgFrag<_Float16> {};
template <> struct WgFrag<e4m3> {
  static constexpr int kPad = 0;
  // CPR = 4 or 8 chunks per row
  template <int CPR> __device__ __forceinline__ static int swz(int r) {
    return CPR == 4 ? (r >> 2) & 3 : (r >> 1) & 7;
  }
  // bytes 0-7 of lane group g: pairs k0 + 8g + e; bytes 8-15: k0 + 32 + 8g + e
  template <int CPR>
  __device__ __forceinline__ static u32x4 load(const char* lds, int pitch, int k0, int c0, int lane) {
    const int j = lane & 15, g = lane >> 4;
    const int ch = c0 + 8 * (j & 1);
    const int r0 = k0 + 8 * g + (j >> 1), r1 = r0 + 32;
    const char* q0 = lds + r0 * pitch + (((ch >> 4) ^ swz<CPR>(r0)) << 4) + (ch & 15);
    const char* q1 = lds + r1 * pitch + (((ch >> 4) ^ swz<CPR>(r1)) << 4) + (ch & 15);
    const wg_v2i32 lo = __builtin_amdgcn_ds_read_tr8_b64_v2i32((WG_LDS wg_v2i32*)(q0));
    const wg_v2i32 hi = __builtin_amdgcn_ds_read_tr8_b64_v2i32((WG_LDS wg_v2i32*)(q1));
    return u32x4{(uint32_t)lo.x, (uint32_t)lo.y, (uint32_t)hi.x, (uint32_t)hi.y};
  }
};

// int8 rows have the byte layout of e4m3 rows; v_mfma_i32_16x16x64_i8 takes
// its 64 reduction bytes per lane group in the order the fp8 pair takes them.
template <> struct WgFrag<int8_t> : WgFrag<e4m3> {};

// Stage geometry of the wgrad kernel for a TM x TN tile: R pairs per stage
// (about 16 KB of rows, a multiple of the MFMA reduction chunk), CPA / CPB
// 16-byte chunks per A / B row, LA / LB chunks per thread and stage.
template <typename T, int TM, int TN>
struct WgGeom {
  static constexpr int CH = Mma<T>::kChunk;
  static constexpr int R0 = 16384 / ((TM + TN) * (int)sizeof(T));
  static constexpr int R = R0 < CH ? CH : (R0 > 128 ? 128 : R0 / CH * CH);
  static constexpr int CPA = TM * (int)sizeof(T) / 16, CPB = TN * (int)sizeof(T) / 16;
  static constexpr int LA = R * CPA / 256, LB = R * CPB / 256;
  static constexpr int PA = TM * (int)sizeof(T) + WgFrag<T>::kPad, PB = TN * (int)sizeof(T) + WgFrag<T>::kPad;
  static_assert(R * CPA % 256 == 0 && R * CPB % 256 == 0, "stage rows must fill the block's loads");
  static_assert(R % CH == 0, "stage must hold whole MFMA chunks");
};

// pieces of offset f's pair list: nb = min(max_blocks, div_up(count, min_rows))
__device__ __forceinline__ int wgrad_pieces(int count, int max_blocks, int min_rows) {
  const int nb = (count + min_rows - 1) / min_rows;
  return nb < max_blocks ? nb : max_blocks;
}

__device__ __forceinline__ int wgrad_total(const WgradArgs& p, int f) {
  return p.fixed_total >= 0 ? p.fixed_total : p.total[f];
}

// Effective mask of every original row: rowmask[argsort[r]] = mask word of
// sorted row r's group & filter.  With it the pair walk below runs in original
// row order (coalesced pair-table reads, ascending output rows) and skips
// exactly the rows the sorted walk would skip.
__global__ __launch_bounds__(256) void wgrad_rowmask_kernel(WgradArgs p) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= p.nrows) return;
  // the argsort entry and the group's mask words load together (unconditional
  // buffer loads; a group past num_mask reads 0), so the store waits one trip
  const int mi = (int)((int64_t)r / p.mask_width);
  const __amdgpu_buffer_rsrc_t rm =
      make_rsrc(p.mask, p.mask != nullptr ? (uint32_t)p.num_mask * (uint32_t)p.mw * 4u : 0u);
  const int i = p.a_rows[r];
  // words in groups of four (dynamic masks carry up to kMaxDynWords)
  for (int w0 = 0; w0 < p.mw; w0 += 4) {
    uint32_t mv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int w = w0 + j;
      mv[j] = __builtin_amdgcn_raw_buffer_load_b32(
          rm, w < p.mw && mi < p.num_mask ? ((uint32_t)mi * (uint32_t)p.mw + (uint32_t)w) * 4u : kBufOOB, 0, 0);
    }
    if ((unsigned)i >= (unsigned)p.nrows) return;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (w0 + j < p.mw)
        p.rowmask[(int64_t)i * p.mw + w0 + j] = p.mask_filter & (p.mask != nullptr ? mv[j] : 0xffffffffu);
  }
}

// Pair-list construction, two passes over [kv] x [blocks of kWgradListRows
// rows].  A thread takes rows i0 + t + 256*k (k < 16), so every load and store
// instruction covers 256 consecutive rows.  Row i holds a pair for offset f
// when pair[f][i] >= 0 and bit f of rowmask[i] is set.
constexpr int kWgradListRows = 4096;

// Loads the 16 rows' pair entries and row masks of one thread (unconditional
// buffer loads; rows past nrows read 0 and are masked by the row test) and
// returns the bits of the rows that hold a pair for offset f.
__device__ __forceinline__ uint32_t wgrad_rows(const WgradArgs& p, int f, int i0, int tid, int (&jv)[16]) {
  const __amdgpu_buffer_rsrc_t rp = make_rsrc(p.pairs + (int64_t)f * p.pair_ld, (uint32_t)p.nrows * 4u);
  // word f / 32 of each row's mask: rowmask is [nrows][mw]
  const __amdgpu_buffer_rsrc_t rm = make_rsrc(p.rowmask + (f >> 5), (uint32_t)p.nrows * (uint32_t)p.mw * 4u);
  uint32_t mv[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int i = i0 + tid + 256 * k;
    const uint32_t off = i < p.nrows ? (uint32_t)i * 4u : kBufOOB;
    jv[k] = (int)__builtin_amdgcn_raw_buffer_load_b32(rp, off, 0, 0);
    const uint32_t moff = i < p.nrows ? (uint32_t)i * (uint32_t)p.mw * 4u : kBufOOB;
    mv[k] = __builtin_amdgcn_raw_buffer_load_b32(rm, moff, 0, 0);
  }
  uint32_t bits = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int i = i0 + tid + 256 * k;
    bits |= (uint32_t)(i < p.nrows && jv[k] >= 0 && ((mv[k] >> (f & 31)) & 1u)) << k;
  }
  return bits;
}

__global__ __launch_bounds__(256) void wgrad_count_kernel(WgradArgs p) {
  const int f = blockIdx.y;
  const int i0 = blockIdx.x * kWgradListRows;
  const int tid = threadIdx.x;
  __shared__ int wsum[4];
  int cnt = 0;
  if ((p.mask_filter >> (f & 31)) & 1u) {
    int jv[16];
    cnt = __popc(wgrad_rows(p, f, i0, tid, jv));
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d);
  if ((tid & 63) == 0) wsum[tid >> 6] = cnt;
  __syncthreads();
  if (tid == 0) p.bcount[f * gridDim.x + blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void wgrad_list_kernel(WgradArgs p) {
  const int f = blockIdx.y;
  const int blk = blockIdx.x;
  const int nblk = gridDim.x;
  const int i0 = blk * kWgradListRows;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (!((p.mask_filter >> (f & 31)) & 1u)) {
    if (blk == 0 && tid == 0) p.total[f] = 0;
    return;
  }
  // base = valid pairs of this offset in earlier blocks; the first 256 counts
  // load as one unconditional buffer load issued with the row loads, so both
  // round trips overlap
  const __amdgpu_buffer_rsrc_t rc = make_rsrc(p.bcount + f * nblk, (uint32_t)blk * 4u);
  int part = (int)__builtin_amdgcn_raw_buffer_load_b32(rc, (uint32_t)tid * 4u, 0, 0);
  int jv[16];
  const uint32_t bits = wgrad_rows(p, f, i0, tid, jv);
  for (int b = tid + 256; b < blk; b += 256) part += p.bcount[f * nblk + b];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d);
  // per-wave counts of every 256-row step k; wpre[k][w] becomes the position
  // of wave w's first pair of step k within the block (steps in order, waves
  // in order inside a step): one 64-lane scan instead of per-entry sums
  __shared__ int wpre[16][4];
  __shared__ int wpart[4];
  uint32_t balc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint64_t bal = __ballot((bits >> k) & 1u);
    balc[k] = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wpre[k][wave] = __popcll(bal);
  }
  if (lane == 0) wpart[wave] = part;
  __syncthreads();
  if (tid < 64) {
    const int c = wpre[tid >> 2][tid & 3];
    int incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (tid >= d) incl += t;
    }
    const int base0 = wpart[0] + wpart[1] + wpart[2] + wpart[3];
    wpre[tid >> 2][tid & 3] = base0 + incl - c;
    if (tid == 63 && blk == nblk - 1) p.total[f] = base0 + incl;
  }
  __syncthreads();
  int32_t* la = p.list_a + (int64_t)f * p.list_ld;
  int32_t* lb = p.list_b + (int64_t)f * p.list_ld;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if ((bits >> k) & 1u) {
      const int off = wpre[k][wave] + (int)balc[k];
      la[off] = i0 + tid + 256 * k;
      lb[off] = jv[k];
    }
  }
}

// Partial products over one piece of offset f's pair list: rows a[list_a] and
// b[list_b], TM x TN output tile (64 or 128 each), FP32 partial per piece.
// Four waves in a 2 x 2 grid.  Every 16-byte load covers a contiguous piece
// of one gathered row (CPA threads per row), and two stages of rows are in
// flight while the MFMAs consume the stage in LDS.
template <typename T, int TM, int TN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void wgrad_kernel(WgradArgs p) {
  using MM = Mma<T>;
  using G = WgGeom<T, TM, TN>;
  constexpr int R = G::R, CH = G::CH, LA = G::LA, LB = G::LB, CPA = G::CPA, CPB = G::CPB;
  constexpr int PA = G::PA, PB = G::PB;
  constexpr int E = 16 / (int)sizeof(T);     // channels per 16-byte chunk
  constexpr int FA = TM / 32, FB = TN / 32;  // 16-wide fragments per wave
  __shared__ __attribute__((aligned(16))) char lds_a[R * PA];
  __shared__ __attribute__((aligned(16))) char lds_b[R * PB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int tiles_n = (p.N + TN - 1) / TN;
  const int tiles = ((p.M + TM - 1) / TM) * tiles_n;
  // Block -> (offset, piece, tile).
  int f = -1, piece = 0, tile = 0;
  if (p.xcd_grid) {
    // kv <= 64, one-dimensional grid.  Workgroups are dealt to the XCDs in
    // turn (block b runs on XCD b % kWgradXcds), and each XCD takes the same
    // eighth of every offset's pieces: offset f's lists run in ascending
    // output row, so an XCD's blocks all gather rows of one eighth of the
    // output rows and their neighbours, and a row needed by several offsets
    // is read from that XCD's L2 after its first use.  Slot s of an XCD is
    // found by a 64-lane scan over the offsets' piece counts (in offset
    // order, tiles innermost).
    const int xcd = (int)blockIdx.x % kWgradXcds, slot = (int)blockIdx.x / kWgradXcds;
    int cnt = 0, lo = 0;
    if (lane < p.kv && ((p.mask_filter >> (lane & 31)) & 1u)) {
      const int nbf = wgrad_pieces(wgrad_total(p, lane), p.max_blocks, p.min_rows);
      lo = nbf * xcd / kWgradXcds;
      cnt = (nbf * (xcd + 1) / kWgradXcds - lo) * tiles;
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    const int rel = slot - (incl - cnt);
    const uint64_t hit = __ballot(cnt > 0 && rel >= 0 && rel < cnt);
    if (!hit) return;
    const int k = __builtin_ctzll(hit);
    const int r = __shfl(rel, k);
    f = k;
    piece = __shfl(lo, k) + r / tiles;
    tile = r % tiles;
  } else {
    // offset-major grid (tiles, kv, max_blocks): larger kernels, where a
    // per-block scan over every offset would cost more than it saves
    f = blockIdx.y;
    piece = blockIdx.z;
    tile = blockIdx.x;
    if (!((p.mask_filter >> (f & 31)) & 1u)) return;
    if (piece >= wgrad_pieces(wgrad_total(p, f), p.max_blocks, p.min_rows)) return;
  }
  const int count = wgrad_total(p, f);
  const int nb = wgrad_pieces(count, p.max_blocks, p.min_rows);
  const int m0 = (tile / tiles_n) * TM;
  const int nn0 = (tile % tiles_n) * TN;
  const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * (TN / 2);
  const int per = (count + nb - 1) / nb;
  const int q0 = piece * per;
  const int nrows_blk = min(per, count - q0);
  const int nst = (nrows_blk + R - 1) / R;
  // this thread's chunk of every A / B row it loads (256 is a multiple of CPA, CPB)
  const int ca = tid % CPA, cb = tid % CPB;
  const bool oka = m0 + ca * E < p.M, okb = nn0 + cb * E < p.N;

  f32x4 acc[FA][FB];
#pragma unroll
  for (int i = 0; i < FA; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // All loads are unconditional buffer loads (out-of-range offsets read 0).
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.a, p.a_bytes);
  const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.b, p.b_bytes);
  const int64_t lofs = (int64_t)f * p.list_ld + q0;
  const __amdgpu_buffer_rsrc_t rLa = make_rsrc(p.list_a + lofs, (uint32_t)nrows_blk * 4u);
  const __amdgpu_buffer_rsrc_t rLb = make_rsrc(p.list_b + lofs, (uint32_t)nrows_blk * 4u);

  auto ld_idx = [&](int s, int (&ia)[LA], int (&ib)[LB]) {
#pragma unroll
    for (int l = 0; l < LA; ++l) {
      const int pos = s * R + (l * 256 + tid) / CPA;
      ia[l] = (int)__builtin_amdgcn_raw_buffer_load_b32(rLa, pos < nrows_blk ? (uint32_t)pos * 4u : kBufOOB, 0, 0);
    }
#pragma unroll
    for (int l = 0; l < LB; ++l) {
      const int pos = s * R + (l * 256 + tid) / CPB;
      ib[l] = (int)__builtin_amdgcn_raw_buffer_load_b32(rLb, pos < nrows_blk ? (uint32_t)pos * 4u : kBufOOB, 0, 0);
    }
  };
  // rows past the piece's end load zeros (out-of-range offsets)
  auto ld_rows = [&](int s, const int (&ia)[LA], const int (&ib)[LB], u32x4 (&va)[LA], u32x4 (&vb)[LB]) {
#pragma unroll
    for (int l = 0; l < LA; ++l) {
      const bool ok = oka && s * R + (l * 256 + tid) / CPA < nrows_blk;
      va[l] = bld16(rA, ok ? (uint32_t)(((int64_t)ia[l] * p.lda + m0 + ca * E) * (int)sizeof(T)) : kBufOOB);
    }
#pragma unroll
    for (int l = 0; l < LB; ++l) {
      const bool ok = okb && s * R + (l * 256 + tid) / CPB < nrows_blk;
      vb[l] = bld16(rB, ok ? (uint32_t)(((int64_t)ib[l] * p.ldb + nn0 + cb * E) * (int)sizeof(T)) : kBufOOB);
    }
  };
  auto st_rows = [&](const u32x4 (&va)[LA], const u32x4 (&vb)[LB]) {
#pragma unroll
    for (int l = 0; l < LA; ++l) {
      const int r = (l * 256 + tid) / CPA;
      *reinterpret_cast<u32x4*>(lds_a + r * PA + ((ca ^ WgFrag<T>::template swz<CPA>(r)) << 4)) = va[l];
    }
#pragma unroll
    for (int l = 0; l < LB; ++l) {
      const int r = (l * 256 + tid) / CPB;
      *reinterpret_cast<u32x4*>(lds_b + r * PB + ((cb ^ WgFrag<T>::template swz<CPB>(r)) << 4)) = vb[l];
    }
  };
  auto compute = [&]() {
#pragma unroll
    for (int kc = 0; kc < R; kc += CH) {
      u32x4 bfr[FB];
#pragma unroll
      for (int j = 0; j < FB; ++j) bfr[j] = WgFrag<T>::template load<CPB>(lds_b, PB, kc, wn + 16 * j, lane);
#pragma unroll
      for (int i = 0; i < FA; ++i) {
        const u32x4 af = WgFrag<T>::template load<CPA>(lds_a, PA, kc, wm + 16 * i, lane);
#pragma unroll
        for (int j = 0; j < FB; ++j) MM::step(acc[i][j], af, bfr[j]);
      }
    }
  };

  // stage s's rows sit in (va0, vb0) for even s and (va1, vb1) for odd s;
  // the stage after next is requested while the MFMAs consume stage s.
  int ia0[LA], ib0[LB], ia1[LA], ib1[LB];
  u32x4 va0[LA], vb0[LB], va1[LA], vb1[LB];
  ld_idx(0, ia0, ib0);
  ld_idx(1, ia1, ib1);
  ld_rows(0, ia0, ib0, va0, vb0);
  ld_rows(1, ia1, ib1, va1, vb1);
  ld_idx(2, ia0, ib0);
  for (int s = 0; s < nst; s += 2) {
    __syncthreads();  // the previous stage's MFMAs are done with the tiles
    st_rows(va0, vb0);
    ld_rows(s + 2, ia0, ib0, va0, vb0);
    ld_idx(s + 3, ia1, ib1);
    __syncthreads();
    compute();
    if (s + 1 >= nst) break;
    __syncthreads();
    st_rows(va1, vb1);
    ld_rows(s + 3, ia1, ib1, va1, vb1);
    ld_idx(s + 4, ia0, ib0);
    __syncthreads();
    compute();
  }

  float* out = p.partial + ((int64_t)f * p.max_blocks + piece) * (int64_t)p.M * p.N;
  const int lr = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int i = 0; i < FA; ++i)
#pragma unroll
    for (int jj = 0; jj < FB; ++jj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int mm = m0 + wm + 16 * i + 4 * lg + e;
        const int nn = nn0 + wn + 16 * jj + lr;
        if (mm < p.M && nn < p.N) out[(int64_t)mm * p.N + nn] = acc[i][jj][e];
      }
}

// ============================================================ split-K tree reduce
// A block reduces 16 * V consecutive output elements.  Thread (q, e) loads V
// consecutive elements of pieces b = q, q + 16, ... (one 16 * V-wide vector
// load per piece, all in flight at once) and sums them; the 16 piece lanes of
// each element are then added in a fixed pairwise tree: the order depends
// only on the piece count, never on scheduling.  V = 4 when the element count
// of an offset is a multiple of 4 (vector loads stay inside one offset).
template <typename OutT, int V, bool INT>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(ReduceArgs p) {
  constexpr int PQ = 16;                     // piece lanes per element
  constexpr int PER = kWgradMaxBlocks / PQ;  // pieces per thread, at most
  // red[q][c * 16 + el]: element el * V + c of piece lane q.  The pitch puts
  // the four piece lanes of one wave on distinct banks (64 banks).
  // INT: the partials are int32 sums, added as unsigned (wrapping) words;
  // every value below is then a bit pattern in a float slot.
  constexpr int RP = V == 4 ? 16 * V + 16 : 16 * V;
  __shared__ float red[PQ][RP];
  const int64_t mn = (int64_t)p.M * p.N;
  const int tid = threadIdx.x;
  const int el = tid & 15, q = tid >> 4;
  const int64_t e = ((int64_t)blockIdx.x * 16 + el) * V;  // first of this thread's V elements
  const bool live = e < (int64_t)p.kv * mn;
  const int fe = live ? (int)(e / mn) : 0;
  const int64_t rem = e - (int64_t)fe * mn;
  const bool on = live && ((p.mask_filter >> (fe & 31)) & 1u);
  const int nb = on ? wgrad_pieces(p.fixed_total >= 0 ? p.fixed_total : p.total[fe], p.max_blocks, p.min_rows) : 0;
  const __amdgpu_buffer_rsrc_t rp = make_rsrc(p.partial + (int64_t)fe * p.max_blocks * mn,
                                              (uint32_t)((int64_t)nb * mn * 4));
  float acc[V];
  if constexpr (V == 4) {
    if constexpr (INT) {
      u32x4 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int b = q + PQ * k;
        v[k] = bld16(rp, b < nb ? (uint32_t)(((int64_t)b * mn + rem) * 4) : kBufOOB);
      }
#pragma unroll
      for (int w = PER / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = __uint_as_float(v[0][c]);
    } else {
      f32x4 v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int b = q + PQ * k;
        const u32x4 x = bld16(rp, b < nb ? (uint32_t)(((int64_t)b * mn + rem) * 4) : kBufOOB);
        v[k] = f32x4{__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z), __uint_as_float(x.w)};
      }
#pragma unroll
      for (int w = PER / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = v[0][c];
    }
  } else {
    if constexpr (INT) {
      uint32_t v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int b = q + PQ * k;
        v[k] = __builtin_amdgcn_raw_buffer_load_b32(rp, b < nb ? (uint32_t)(((int64_t)b * mn + rem) * 4) : kBufOOB,
                                                    0, 0);
      }
#pragma unroll
      for (int w = PER / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
      acc[0] = __uint_as_float(v[0]);
    } else {
      float v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int b = q + PQ * k;
        v[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
            rp, b < nb ? (uint32_t)(((int64_t)b * mn + rem) * 4) : kBufOOB, 0, 0));
      }
#pragma unroll
      for (int w = PER / 2; w >= 1; w >>= 1)
#pragma unroll
        for (int i = 0; i < w; ++i) v[i] = v[2 * i] + v[2 * i + 1];
      acc[0] = v[0];
    }
  }
#pragma unroll
  for (int c = 0; c < V; ++c) red[q][c * 16 + el] = acc[c];
  __syncthreads();
  // 16 * V elements, 256 threads: thread tid finishes element tid (if any)
  if (tid >= 16 * V) return;
  const int64_t eo = (int64_t)blockIdx.x * 16 * V + tid;
  if (eo >= (int64_t)p.kv * mn) return;
  const int fo = (int)(eo / mn);
  if (!((p.mask_filter >> (fo & 31)) & 1u)) return;
  const int64_t ro = eo - (int64_t)fo * mn;
  const int mm = (int)(ro / p.N);
  const int nn = (int)(ro - (int64_t)mm * p.N);
  const int64_t oi = (int64_t)mm * p.o_ms + (int64_t)fo * p.o_fs + nn;
  const OutT* sp = reinterpret_cast<const OutT*>(p.src);
  // a bias broadcasts over the GEMM rows and is indexed by the GEMM column,
  // which for the [K, kv * C] view of the KRSC weight gradient is f * C + c
  const int64_t si = p.src_is_bias ? (int64_t)fo * p.N + nn : (int64_t)mm * p.s_ms + (int64_t)fo * p.s_fs + nn;
  if constexpr (INT) {
    uint32_t t[PQ];
#pragma unroll
    for (int i = 0; i < PQ; ++i) t[i] = __float_as_uint(red[i][(tid % V) * 16 + tid / V]);
#pragma unroll
    for (int w = PQ / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int i = 0; i < w; ++i) t[i] = t[2 * i] + t[2 * i + 1];
    const int32_t src = p.ibeta != 0 ? to_i32(sp[si]) : 0;
    reinterpret_cast<OutT*>(p.out)[oi] = int_out<OutT>(int_epilogue(p, (int32_t)t[0], src));
  } else {
    float t[PQ];
#pragma unroll
    for (int i = 0; i < PQ; ++i) t[i] = red[i][(tid % V) * 16 + tid / V];
#pragma unroll
    for (int w = PQ / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int i = 0; i < w; ++i) t[i] = t[2 * i] + t[2 * i + 1];
    float r = p.alpha * t[0];
    if (p.beta != 0.f) r += p.beta * to_f32(sp[si]);
    r = apply_act(r, p.act_type, p.act_alpha);
    reinterpret_cast<OutT*>(p.out)[oi] = from_f32<OutT>(r);
  }
}

// ============================================================ helpers
// out[f][c][r] = in[r*in_row_stride + f*Cc + c]: one 32x32 tile of one f per
// block, staged through LDS so both the read and the write are coalesced.
template <typename U>
__global__ __launch_bounds__(256) void transpose_kernel(const U* in, U* out, int R, int Cc,
                                                        int64_t in_row_stride) {
  __shared__ U tile[32][33];
  const int f = blockIdx.z;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
  for (int i = 0; i < 32; i += 8) {
    const int r = r0 + ty + i, c = c0 + tx;
    if (r < R && c < Cc) tile[ty + i][tx] = in[(int64_t)r * in_row_stride + (int64_t)f * Cc + c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 32; i += 8) {
    const int c = c0 + ty + i, r = r0 + tx;
    if (r < R && c < Cc) out[((int64_t)f * Cc + c) * R + r] = tile[tx][ty + i];
  }
}

__global__ void zero_kernel(uint8_t* p, int64_t nbytes) {
  const int64_t nwords = nbytes >> 2;
  uint32_t* w = reinterpret_cast<uint32_t*>(p);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0u;
  const int64_t t = (nwords << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nbytes) p[t] = 0u;
}

// ============================================================ launchers
int rowgather_block_n(int nout) { return nout <= 64 ? 64 : 128; }

bool rowgather_supported(int in_dtype, int out_dtype) {
  switch (in_dtype) {
    case kF32: return out_dtype == kF32;
    case kF16: return out_dtype == kF16 || out_dtype == kF32;
    case kBF16: return out_dtype == kBF16 || out_dtype == kF32;
    case kE4M3: return out_dtype == kF16 || out_dtype == kBF16 || out_dtype == kF32;
    case kI8: return out_dtype == kI8;
    default: return false;
  }
}

hipError_t launch_rowgather(int in_dtype, int out_dtype, int block_n, const RowGatherArgs& a, hipStream_t s) {
  switch (in_dtype) {
    case kF32: return launch_rowgather_f32(out_dtype, block_n, a, s);
    case kF16: return launch_rowgather_f16(out_dtype, block_n, a, s);
    case kBF16: return launch_rowgather_bf16(out_dtype, block_n, a, s);
    case kE4M3: return launch_rowgather_e4m3(out_dtype, block_n, a, s);
    case kI8: return launch_rowgather_i8(out_dtype, block_n, a, s);
    default: return hipErrorInvalidValue;
  }
}

const void*& last_kernel() {
  static thread_local const void* k = nullptr;
  return k;
}

bool gemm_rows_supported(int in_dtype, int out_dtype) {
  switch (in_dtype) {
    case kF32: return out_dtype == kF32;
    case kF16: return out_dtype == kF16;
    case kBF16: return out_dtype == kBF16;
    case kI8: return out_dtype == kI8 || out_dtype == kI32;
    default: return false;
  }
}

hipError_t launch_gemm_rows(int in_dtype, int out_dtype, const GemmRowsArgs& a, hipStream_t s) {
  switch (in_dtype) {
    case kF32: return launch_gemm_rows_f32(out_dtype, a, s);
    case kF16: return launch_gemm_rows_f16(out_dtype, a, s);
    case kBF16: return launch_gemm_rows_bf16(out_dtype, a, s);
    case kI8: return launch_gemm_rows_i8(out_dtype, a, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
static hipError_t launch_wgrad_t(int tm, int tn, dim3 grid, const WgradArgs& a, hipStream_t s) {
  dim3 block(256, 1, 1);
  if (tm == 64 && tn == 64) CS_LAUNCH((wgrad_kernel<T, 64, 64>), grid, block, 0, s, a);
  else if (tm == 64) CS_LAUNCH((wgrad_kernel<T, 64, 128>), grid, block, 0, s, a);
  else if (tn == 64) CS_LAUNCH((wgrad_kernel<T, 128, 64>), grid, block, 0, s, a);
  else CS_LAUNCH((wgrad_kernel<T, 128, 128>), grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_wgrad(int in_dtype, const WgradArgs& a, hipStream_t s) {
  if (a.max_blocks < 1 || a.max_blocks > kWgradMaxBlocks || a.min_rows < 1) return hipErrorInvalidValue;
  hipError_t e;
  if (a.fixed_total < 0) {
    // conv form: effective row masks, then the per-offset pair lists
    if (a.nrows > 0) {
      hipLaunchKernelGGL(wgrad_rowmask_kernel, dim3((a.nrows + 255) / 256), dim3(256), 0, s, a);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int nblk = std::max(1, (a.nrows + kWgradListRows - 1) / kWgradListRows);
    hipLaunchKernelGGL(wgrad_count_kernel, dim3(nblk, a.kv), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(wgrad_list_kernel, dim3(nblk, a.kv), dim3(256), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // 128-wide tiles where a dimension exceeds 64: each row is gathered once per
  // tile.  A single list (ShuffleAB) with few pieces keeps 64-wide tiles, so
  // the grid has four times the blocks (rows read twice come from cache).
  const int pieces = a.fixed_total >= 0
                         ? std::min(a.max_blocks, std::max(1, (a.fixed_total + a.min_rows - 1) / a.min_rows))
                         : a.max_blocks;
  const bool few = a.fixed_total >= 0 && a.kv == 1 && pieces < 256;
  const int tm = a.M > 64 && !few ? 128 : 64, tn = a.N > 64 && !few ? 128 : 64;
  const int tiles = ((a.M + tm - 1) / tm) * ((a.N + tn - 1) / tn);
  // kv <= 64: one-dimensional grid grouped per XCD, where an XCD takes at
  // most ceil(max_blocks / kWgradXcds) pieces of each offset per tile (blocks
  // past its share return at once); else (tiles, kv, pieces)
  // conv form: an offset's list holds at most one pair per output row, so no
  // offset has more than ceil(nrows / min_rows) pieces; the grid stops there
  // (the partition, which uses max_blocks, is unchanged)
  int mb = a.max_blocks;
  if (a.fixed_total < 0)
    mb = (int)std::min<int64_t>(mb, std::max<int64_t>(1, ((int64_t)a.nrows + a.min_rows - 1) / a.min_rows));
  WgradArgs w = a;
  w.xcd_grid = a.kv <= 64;
  dim3 grid = w.xcd_grid ? dim3((unsigned)(kWgradXcds * tiles * a.kv * ((mb + kWgradXcds - 1) / kWgradXcds)), 1, 1)
                         : dim3((unsigned)tiles, (unsigned)a.kv, (unsigned)mb);
  dim3 block(256, 1, 1);
  if (tiles == 0) return hipSuccess;
  switch (in_dtype) {
    case kF32: return launch_wgrad_t<float>(tm, tn, grid, w, s);
    case kF16: return launch_wgrad_t<_Float16>(tm, tn, grid, w, s);
    case kBF16: return launch_wgrad_t<bf16>(tm, tn, grid, w, s);
    case kE4M3: return launch_wgrad_t<e4m3>(tm, tn, grid, w, s);
    case kI8: return launch_wgrad_t<int8_t>(tm, tn, grid, w, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename OutT, bool INT>
static hipError_t launch_reduce_t(const ReduceArgs& a, hipStream_t s) {
  const int64_t mn = (int64_t)a.M * a.N;
  const int64_t total = (int64_t)a.kv * mn;
  if (total == 0) return hipSuccess;
  if (a.max_blocks < 1 || a.max_blocks > kWgradMaxBlocks) return hipErrorInvalidValue;
  if (mn % 4 == 0)
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 4, INT>), dim3((unsigned)((total + 63) / 64)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((splitk_reduce_kernel<OutT, 1, INT>), dim3((unsigned)((total + 15) / 16)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_reduce(int out_dtype, const ReduceArgs& a, hipStream_t s) {
  if (a.int_compute) {
    switch (out_dtype) {
      case kI8: return launch_reduce_t<int8_t, true>(a, s);
      case kI32: return launch_reduce_t<int32_t, true>(a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (out_dtype) {
    case kF32: return launch_reduce_t<float, false>(a, s);
    case kF16: return launch_reduce_t<_Float16, false>(a, s);
    case kBF16: return launch_reduce_t<bf16, false>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_transpose(const void* in, void* out, int64_t R, int64_t F, int64_t Cc,
                            int64_t in_row_stride, int elem_bytes, hipStream_t s) {
  if (R * F * Cc == 0) return hipSuccess;
  if (R > INT32_MAX || Cc > INT32_MAX || F > 65535) return hipErrorInvalidValue;
  dim3 grid((unsigned)((Cc + 31) / 32), (unsigned)((R + 31) / 32), (unsigned)F), block(256);
  switch (elem_bytes) {
    case 1:
      hipLaunchKernelGGL(transpose_kernel<uint8_t>, grid, block, 0, s, reinterpret_cast<const uint8_t*>(in),
                         reinterpret_cast<uint8_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
    case 2:
      hipLaunchKernelGGL(transpose_kernel<uint16_t>, grid, block, 0, s, reinterpret_cast<const uint16_t*>(in),
                         reinterpret_cast<uint16_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
    case 4:
      hipLaunchKernelGGL(transpose_kernel<uint32_t>, grid, block, 0, s, reinterpret_cast<const uint32_t*>(in),
                         reinterpret_cast<uint32_t*>(out), (int)R, (int)Cc, in_row_stride);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_zero(void* p, int64_t nbytes, hipStream_t s) {
  if (nbytes <= 0) return hipSuccess;
  // the buffer base must be 4-byte aligned (device allocations are)
  const int64_t nwords = (nbytes + 3) / 4;
  int blocks = (int)std::min<int64_t>((nwords + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint8_t*>(p), nbytes);
  return hipGetLastError();
}

}  // namespace cs
