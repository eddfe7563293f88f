// Row-gather implicit GEMM kernel (sparse conv forward / dgrad, ShuffleAC and
// dense GEMM).
//
// A block owns 64 sorted rows (four 16-row mask tiles) and 64 output columns.
// Its four waves OR the four tile masks into the same scalar union and split
// the union's kernel offsets by offset index (wave w takes f = w, w+4, ...),
// so the longest dependent chain of a block is about a quarter of its offsets
// and each row's summation order does not depend on its tile's other rows.  Each
// wave accumulates all 64x64 outputs of its offsets in MFMA registers:
// gathered rows and weight rows are loaded straight into operand registers
// (16 B per lane) with raw buffer loads whose out-of-range offset reads zero,
// used for rows without a neighbour (index -1) and tiles without the offset,
// so no branch or select waits on the load stream.  Offsets are taken two at a
// time (one memory round trip per pair) and the next pair's indices load while
// the MFMAs run; tiles whose own mask lacks an offset skip its MFMAs.  At the
// end the four partial tiles are summed through LDS in wave order (fixed,
// deterministic) and each wave writes 16 rows with whole-row coalesced stores
// through the epilogue.
// Instantiated per input dtype in rg_<dtype>.hip.
#pragma once
#include <climits>
#include "cs_device.h"
#include "cs_kernels.h"
#include "cs_rowgather_stream.h"
#include "cs_rowgather_ws.h"

namespace cs {

// Reduction chunks held per offset and pass.
template <typename T> struct RgChunks { static constexpr int value = 2; };

constexpr int kRgRows = 64;       // rows per block (4 mask tiles)
constexpr int kRgRedPitch = 68;   // LDS floats per row of a partial tile

// HAS_PAIRS: gather through a pair table (else the identity / a_inds-free GEMM).
// FULL: kred is a multiple of the per-pass reduction width and nout a multiple
// of 64, so no lane needs a reduction- or column-range predicate.
// MW: mask words per window of offsets (1, or kMaxMaskWords for dynamic
// masks, kv > 32, walked window by window).
template <typename T, typename OutT, bool HAS_PAIRS, bool FULL, int MW>
__global__ __launch_bounds__(256) void rowgather_kernel(RowGatherArgs p) {
  using MM = Mma<T>;
  constexpr int E = MM::kElems;
  constexpr int CH = MM::kChunk;
  constexpr int KCH = RgChunks<T>::value;
  constexpr int NB = 4;
  __shared__ float red[4][kRgRows][kRgRedPitch];
  __shared__ uint32_t tmw[MW > 1 ? 4 : 1][MW > 1 ? 4 : 1][MW > 1 ? kMaxDynWords : 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // Rows are sorted by mask, and the blocks late in that order carry the most
  // offsets (the high offsets are the sort's most significant bits): blocks
  // take the row groups from the last one back, so the longest start first and
  // every XCD (workgroups are dealt to the 8 XCDs in turn) gets an even share.
  const int tile_blk = (int)gridDim.x - 1 - (int)blockIdx.x;
  const int row0 = tile_blk * kRgRows;
  if (row0 >= p.m) return;  // whole block leaves together
  const int n0 = blockIdx.y * (NB * 16);
  const int lr = lane & 15;
  const int lg = lane >> 4;

  const int myrow = row0 + lane;
  const bool rvalid = myrow < p.m;
  int mycol = 0, myout = 0;
  if (rvalid) {
    mycol = p.pair_col ? p.pair_col[myrow] : myrow;
    myout = p.out_row ? p.out_row[myrow] : myrow;
  }
  // OR of the 16 row masks of each tile (lanes 16s..16s+15 hold tile s)
  uint32_t sm[4][MW];
  uint32_t live = 0;
  if constexpr (MW == 1) {
    uint32_t g = 0u;
    if (rvalid) g = p.mask ? p.mask[myrow] : 0xffffffffu;
    g |= __shfl_xor(g, 1);
    g |= __shfl_xor(g, 2);
    g |= __shfl_xor(g, 4);
    g |= __shfl_xor(g, 8);
#pragma unroll
    for (int s = 0; s < 4; ++s) sm[s][0] = __builtin_amdgcn_readlane(g, 16 * s) & p.mask_filter & kv_word(p.kv, 0);
    if (wave == 0 && p.mask_out != nullptr && lane < 4 && row0 + 16 * lane < p.m) {
      const uint32_t v = lane == 0 ? sm[0][0] : lane == 1 ? sm[1][0] : lane == 2 ? sm[2][0] : sm[3][0];
      p.mask_out[row0 / kMaskTile + lane] = v;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (row0 + 16 * s < p.m && (!p.skip_empty || sm[s][0] != 0u)) live |= 1u << s;
    if (p.reverse_mask) {
#pragma unroll
      for (int s = 0; s < 4; ++s) reverse_kv(sm[s], p.kv);
    }
  } else {
    // dynamic masks: the four tiles' OR words, all of them, in this wave's
    // LDS area (every wave holds all 64 rows, so no barrier is needed); the
    // offsets are then walked in windows of MW words, in kernel order
    for (int q0 = 0; q0 < p.mw; q0 += 8) {
      uint32_t g[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        g[j] = 0u;
        if (rvalid && q0 + j < p.mw) g[j] = p.mask ? p.mask[(int64_t)myrow * p.mw + q0 + j] : 0xffffffffu;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        uint32_t v = g[j];
        v |= __shfl_xor(v, 1);
        v |= __shfl_xor(v, 2);
        v |= __shfl_xor(v, 4);
        v |= __shfl_xor(v, 8);
        if (lr == 0 && q0 + j < p.mw) tmw[wave][lg][q0 + j] = v & p.mask_filter & kv_word(p.kv, q0 + j);
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      uint32_t any = 0u;
      for (int q = 0; q < p.mw; ++q) any |= tmw[wave][s][q];
      if (row0 + 16 * s < p.m && (!p.skip_empty || __builtin_amdgcn_readfirstlane(any) != 0u)) live |= 1u << s;
      if (wave == 0 && p.mask_out != nullptr && row0 + 16 * s < p.m)
        for (int q = lane; q < p.mw; q += 64) p.mask_out[(int64_t)(row0 / kMaskTile + s) * p.mw + q] = tmw[0][s][q];
    }
  }

  int cols[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) cols[s] = __shfl(mycol, 16 * s + lr);

  f32x4 acc[4][NB];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[s][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  // pair indices of offset f for the rows this lane feeds (-1: no neighbour).
  // The loaded value is not touched here, so the load stays in flight.
  int base = 0;  // first offset of the current window (dynamic masks)
  auto load_idx = [&](int f, int (&idx)[4]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) idx[s] = HAS_PAIRS ? p.pairs[(int64_t)(base + f) * p.pair_ld + cols[s]] : cols[s];
  };
  // operand fragments of offset f (f < 0: none) for reduction block kb; rows
  // with index -1, tiles without f, columns past nout and reduction elements
  // past kred get an out-of-range offset and read zeros
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.a, p.a_bytes);
  const __amdgpu_buffer_rsrc_t rb = make_rsrc(p.b, p.b_bytes);
  auto load_frags = [&](int f, const int (&idx)[4], int kb, u32x4 (&af)[4][KCH], u32x4 (&bfr)[KCH][NB]) {
    const bool fvalid = f >= 0;
    const int64_t fb = (int64_t)(fvalid ? base + f : 0) * p.b_fs;
#pragma unroll
    for (int kc = 0; kc < KCH; ++kc) {
      const int k = kb + kc * CH + lg * E;
      const bool kin = fvalid && (FULL || k < p.kred);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int n = n0 + nb * 16 + lr;
        const uint32_t off = (kin && (FULL || n < p.nout))
                                 ? (uint32_t)((fb + (int64_t)n * p.b_ns + k) * (int64_t)sizeof(T)) : kBufOOB;
        bfr[kc][nb] = bld16(rb, off);
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool on = kin && mask_bit(sm[s], f) != 0u && idx[s] >= 0;
        const uint32_t off = on ? (uint32_t)(((int64_t)idx[s] * p.lda + k) * (int64_t)sizeof(T)) : kBufOOB;
        af[s][kc] = bld16(ra, off);
      }
    }
  };
  auto compute = [&](int f, const u32x4 (&af)[4][KCH], const u32x4 (&bfr)[KCH][NB]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (mask_bit(sm[s], f)) {
#pragma unroll
        for (int kc = 0; kc < KCH; ++kc)
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) MM::step(acc[s][nb], af[s][kc], bfr[kc][nb]);
      }
    }
  };

  // This wave's offsets, two at a time: a pair's fragment loads form one
  // memory round trip, and the next pair's indices load while the MFMAs run.
  auto window = [&]() {
  // This wave's share: offsets f with f % 4 == wave.  The split depends on the
  // offset index only, never on which offsets the tile's union holds, so every
  // row's partial sums (and hence its bits) are the same whatever rows share
  // its tile: a cloud gives identical rows alone and inside a batch.
  uint32_t mine[MW];
  uint32_t mine_any = 0u;
#pragma unroll
  for (int w = 0; w < MW; ++w) {
    mine[w] = (sm[0][w] | sm[1][w] | sm[2][w] | sm[3][w]) & (0x11111111u << wave);
    mine_any |= mine[w];
  }
  for (int kb = 0; kb < p.kred; kb += KCH * CH) {
    if (!mine_any) break;
    uint32_t rem[MW];
#pragma unroll
    for (int w = 0; w < MW; ++w) rem[w] = mine[w];
    auto pop = [&]() -> int {  // lowest remaining offset, in ascending order
      int f = -1;
#pragma unroll
      for (int w = MW - 1; w >= 0; --w) {
        if (rem[w]) f = 32 * w + __builtin_ctz(rem[w]);
      }
      if (f >= 0) {
#pragma unroll
        for (int w = 0; w < MW; ++w)
          if ((f >> 5) == w) rem[w] &= rem[w] - 1u;
      }
      return f;
    };
    int fa = pop(), fb = pop();
    int Ia[4], Ib[4];
    load_idx(fa, Ia);
    load_idx(fb < 0 ? fa : fb, Ib);
    while (true) {
      u32x4 Aa[4][KCH], Ba[KCH][NB], Ab[4][KCH], Bb[KCH][NB];
      load_frags(fa, Ia, kb, Aa, Ba);
      load_frags(fb, Ib, kb, Ab, Bb);
      const int fc = pop(), fd = pop();
      if (fc >= 0) {
        load_idx(fc, Ia);
        load_idx(fd < 0 ? fc : fd, Ib);
      }
      compute(fa, Aa, Ba);
      if (fb >= 0) compute(fb, Ab, Bb);
      if (fc < 0) break;
      fa = fc;
      fb = fd;
    }
  }
  };
  if constexpr (MW == 1) {
    window();
  } else {
    const bool rev = p.reverse_mask != 0;
    for (int win = 0; 32 * MW * win < p.kv; ++win) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        auto word = [&](int q) -> uint32_t { return q >= 0 && q < p.mw ? tmw[wave][s][q] : 0u; };
#pragma unroll
        for (int i = 0; i < MW; ++i) sm[s][i] = __builtin_amdgcn_readfirstlane(kernel_word(word, MW * win + i, p.kv, rev));
      }
      base = 32 * MW * win;
      window();
    }
  }

  // partial tiles -> LDS; accumulator element i of block (s, nb) is row
  // 16s + 4lg + i, column 16nb + lr
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[wave][16 * s + 4 * lg + i][16 * nb + lr] = acc[s][nb][i];
  __syncthreads();

  // wave w finishes tile w: fixed-order sum of the four partials, epilogue,
  // one 64-column row per store instruction
  const int s = wave;
  if (!((live >> s) & 1u)) return;
  OutT* C = reinterpret_cast<OutT*>(p.c);
  const OutT* D = reinterpret_cast<const OutT*>(p.d);
  const int n = n0 + lane;
  const bool ncol = n < p.nout;
  const float scl = (p.scale && ncol) ? p.scale[n] : 1.f;
  const float bia = (IsInt8<T>::value && ncol) ? p.bias_f[n] : 0.f;
  const float bias = (p.beta != 0.f && p.d_is_bias && ncol) ? to_f32(D[n]) : 0.f;
#pragma unroll 4
  for (int i = 0; i < 16; ++i) {
    const int rr = 16 * s + i;
    const int orow = __shfl(myout, rr);
    if (row0 + rr < p.m && ncol) {
      float v;
      if constexpr (IsInt8<T>::value) {
        const int32_t acc = ((__float_as_int(red[0][rr][lane]) + __float_as_int(red[1][rr][lane])) +
                             __float_as_int(red[2][rr][lane])) + __float_as_int(red[3][rr][lane]);
        v = i8_epilogue(p, acc, scl, bia, n, orow);
      } else {
        v = ((red[0][rr][lane] + red[1][rr][lane]) + red[2][rr][lane]) + red[3][rr][lane];
        v *= scl;
        v *= p.alpha;
        if (p.beta != 0.f) v += p.beta * (p.d_is_bias ? bias : to_f32(D[(int64_t)orow * p.ldd + n]));
        v = apply_act(v, p.act_type, p.act_alpha);
      }
      C[(int64_t)orow * p.ldc + n] = from_f32<OutT>(v);
    }
  }
}

// Kernel choice.  block_n (the algo's tile_shape[1]) of 16 / 32 / 64 caps the
// weight-stationary kernel's column slice; with 64 that kernel runs when it
// needs at most two column slices (four for f32) and the streamed-weights
// kernel when it would need more (every slice gathers the input rows again);
// 128 asks for the streamed-weights kernel.  The block kernel covers the rest.
template <typename T>
static int ws_slices(const RowGatherArgs& a) {
  int ncol = std::min(ws_ncol<T>(a.kv, a.kred), 64);
  if (ncol == 48) ncol = 32;
  if ((a.kred + Mma<T>::kChunk - 1) / Mma<T>::kChunk == 4 && ncol > 32) ncol = 32;  // as launch_ws
  return ncol > 0 ? (a.nout + ncol - 1) / ncol : INT_MAX;
}

template <typename T, typename OutT, int MW>
static hipError_t launch_block(const RowGatherArgs& a, hipStream_t s) {
  dim3 grid((a.m + kRgRows - 1) / kRgRows, (a.nout + 63) / 64, 1);
  dim3 block(256, 1, 1);
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  constexpr int pass = RgChunks<T>::value * Mma<T>::kChunk;
  const bool full = a.kred % pass == 0 && a.nout % 64 == 0;
  if (a.pairs) {
    if (full) CS_LAUNCH((rowgather_kernel<T, OutT, true, true, MW>), grid, block, 0, s, a);
    else CS_LAUNCH((rowgather_kernel<T, OutT, true, false, MW>), grid, block, 0, s, a);
  } else if constexpr (MW == 1) {
    if (full) CS_LAUNCH((rowgather_kernel<T, OutT, false, true, 1>), grid, block, 0, s, a);
    else CS_LAUNCH((rowgather_kernel<T, OutT, false, false, 1>), grid, block, 0, s, a);
  } else {
    return hipErrorInvalidValue;  // multi-word masks come with a pair table
  }
  return hipGetLastError();
}

// B_f n-major copy of k-major weights into b_scratch, then the launch on it.
template <typename T, typename OutT>
static hipError_t launch_rg(const RowGatherArgs& a, int block_n, hipStream_t s);

template <typename T, typename OutT>
static hipError_t launch_transposed(const RowGatherArgs& a, int block_n, hipStream_t s) {
  // the other kernels read B n-major: transpose [kred][kv][nout] -> [kv][nout][kred]
  if (a.b_scratch == nullptr) return hipErrorInvalidValue;
  const hipError_t et = launch_transpose(a.b, a.b_scratch, a.kred, a.kv, a.nout, a.b_ks, (int)sizeof(T), s);
  if (et != hipSuccess) return et;
  RowGatherArgs a2 = a;
  a2.b = a.b_scratch;
  a2.b_kmajor = 0;
  a2.b_fs = (int64_t)a.nout * a.kred;
  a2.b_ns = a.kred;
  a2.b_bytes = (uint32_t)((int64_t)a.kv * a.nout * a.kred * (int64_t)sizeof(T));
  return launch_rg<T, OutT>(a2, block_n, s);
}

template <typename T, typename OutT>
static hipError_t launch_rg(const RowGatherArgs& a, int block_n, hipStream_t s) {
  if (block_n != 16 && block_n != 32 && block_n != 64 && block_n != 128) return hipErrorInvalidValue;
  if (a.mw < 1 || a.mw > kMaxDynWords) return hipErrorInvalidValue;
  bool used = false;
  hipError_t e;
  if (a.mw > 1) {
    // dynamic masks (kv > 32): streamed weights where they apply, else the block kernel
    if (a.b_kmajor) return launch_transposed<T, OutT>(a, block_n, s);
    e = launch_stream<T, OutT, kMaxMaskWords>(a, s, &used);
    if (used || e != hipSuccess) return e;
    return launch_block<T, OutT, kMaxMaskWords>(a, s);
  }
  // measured on MI355X: two slices beat streaming for 1- and 2-byte types, four for f32
  const bool prefer_stream = block_n == 128 || (block_n == 64 && ws_slices<T>(a) > (sizeof(T) == 4 ? 4 : 2));
  if (!prefer_stream) {
    e = launch_ws<T, OutT>(a, block_n, s, &used);
    if (used || e != hipSuccess) return e;
  }
  if (a.b_kmajor) return launch_transposed<T, OutT>(a, block_n, s);
  e = launch_stream<T, OutT, 1>(a, s, &used);
  if (used || e != hipSuccess) return e;
  if (prefer_stream) {
    e = launch_ws<T, OutT>(a, 64, s, &used);
    if (used || e != hipSuccess) return e;
  }
  return launch_block<T, OutT, 1>(a, s);
}

}  // namespace cs
