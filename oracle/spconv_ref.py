"""FP64 restatement of ConvMainUnitTest.implicit_gemm2 (mask-sparse) semantics,
plus the reference's CPU gather-GEMM-scatter path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Notation: X [N_in, C] features, W [K, kv, C] weights (KRSC viewed as
``weight.view([K, -1, C])``, cumm/conv/main_real_spconv.py:255), pair [kv, N]
int32 with -1 for no neighbour, argsort [N] int32, mask [N] uint32 in sorted
order, tile_m = the kernel's ``tile_shape[0]``.

Pinning (tests/test_oracle.py):
  * conv_forward      -- bruteforce.conv_forward_coords on tiny grids (independent
                         neighbour search), the reference's own compact-pair
                         algorithm (cpu_fwd below), centre-only tables reducing to
                         a dense GEMM, all -1 tables giving zeros, tile-mask
                         granularity on hand-built split masks.
  * conv_dgrad        -- adjoint identity <fwd(X), dO> == <X, dgrad(dO)> on the
                         same pairs, cpu_dgrad, reverse-mask equivalence for subm.
  * conv_wgrad        -- adjoint identity <fwd(X;W), dO> == <W, wgrad(X, dO)>,
                         cpu_wgrad, reduced-mask skipping on hand-built masks.
  * int8_inference_epilogue -- hand-computed known answers for rounding ties
                         (half to even), saturation at -128 / 127, bias and
                         output_add terms, activation before clamping; with
                         scale 1 and bias 0 the int8 conv equals the integer
                         brute-force sum clipped to [-128, 127].
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

ALL = 0xFFFFFFFF


def _act(v: np.ndarray, act_type: int, act_alpha: float) -> np.ndarray:
    """UnaryActivation (cumm/gemm/gemmmath/__init__.py:152-222): None, ReLU,
    Sigmoid, LeakyReLU."""
    if act_type == 1:
        return np.maximum(v, 0.0)
    if act_type == 2:
        return 1.0 / (1.0 + np.exp(-v))
    if act_type == 3:
        return np.where(v >= 0.0, v, v * act_alpha)
    return v


def epilogue(acc: np.ndarray, alpha: float, beta: float, src: Optional[np.ndarray],
             act_type: int = 0, act_alpha: float = 0.0) -> np.ndarray:
    """LinearCombination (cumm/gemm/output_op/linear.py:108-195):
    ``act(alpha * acc + beta * src)``.  The source is not read when beta == 0."""
    v = alpha * acc
    if beta != 0.0 and src is not None:
        v = v + beta * src
    return _act(v, act_type, act_alpha)


def int8_inference_epilogue(acc, scale, bias, beta: float = 0.0, source=None, act_type: int = 0,
                            act_alpha: float = 0.0) -> np.ndarray:
    """Int8Inference output op without pre-scaled bias
    (cumm/gemm/output_op/linear.py:198-392, used when ``is_int8_inference``,
    cumm/gemm/algospec/turing.py:480-500): in float32,
    ``v = scale * float(acc) + bias`` (one fused multiply-add), then
    ``v = beta * source + v`` when beta != 0, then the activation, then clamp to
    [-128, 127] and round to nearest, ties to even, into int8.  alpha is not
    used.  acc is the exact int32 accumulator [N, K]; scale / bias are per
    output channel; source is output_add (or the output) [N, K] int8.

    Each fused multiply-add is evaluated exactly in float64 (the product of two
    float32 values is exact there) and rounded once to float32; this equals a
    float32 FMA unless the float64 sum rounds onto a float32 half-way point,
    which needs a 29-bit coincidence and does not occur in the tests' data.
    """
    f32 = np.float32
    a = np.asarray(acc, np.int64).astype(f32).astype(np.float64)
    sc = np.asarray(scale, f32).astype(np.float64)[None, :a.shape[1]]
    bi = np.asarray(bias, f32).astype(np.float64)[None, :a.shape[1]]
    v = (sc * a + bi).astype(f32)
    if beta != 0.0:
        src = np.asarray(source, np.int64).astype(np.float64)
        v = (np.float64(f32(beta)) * src + v.astype(np.float64)).astype(f32)
    v = _act(v.astype(np.float64), act_type, float(f32(act_alpha))).astype(f32)
    v = np.clip(v, f32(-128), f32(127))
    return np.rint(v).astype(np.int8)


def tile_masks(mask_sorted: np.ndarray, n: int, tile_m: int, mask_filter: int = ALL) -> np.ndarray:
    """Per-tile OR of the sorted row masks AND mask_filter
    (cumm/conv/kernel.py:843-866; rows past N contribute 0).  Masks are [N]
    (kv <= 32) or [N, W] words for dynamic masks, where the same 32-bit filter
    applies to every word (MaskIGemmIteratorMaskLoaderDynamic::load_mask,
    cumm/gemm/blockmma/mma.py:168-199); the result has the input's rank."""
    t = (n + tile_m - 1) // tile_m
    m = np.asarray(mask_sorted)
    words = m.reshape(m.shape[0], -1)[:n].astype(np.uint64)
    pad = np.zeros((t * tile_m, words.shape[1]), dtype=np.uint64)
    pad[:n] = words
    red = np.bitwise_or.reduce(pad.reshape(t, tile_m, -1), axis=1) & np.uint64(mask_filter)
    return red[:, 0] if m.ndim == 1 else red


def mask_bits(masks: np.ndarray, kv: int) -> np.ndarray:
    """[R] or [R, W] mask words -> bool [R, kv]: offset f is bit f % 32 of word f // 32."""
    m = np.asarray(masks).astype(np.uint64)
    m = m.reshape(m.shape[0], -1)
    f = np.arange(kv)
    return ((m[:, f // 32] >> (f % 32).astype(np.uint64)) & np.uint64(1)).astype(bool)


def _pack(bits: np.ndarray, like: np.ndarray) -> np.ndarray:
    """bool [R, kv] -> words shaped like ``like`` ([R] or [R, W])."""
    w = 1 if like.ndim == 1 else like.shape[1]
    out = np.zeros((bits.shape[0], w), dtype=np.uint64)
    for f in range(bits.shape[1]):
        out[:, f // 32] |= bits[:, f].astype(np.uint64) << np.uint64(f % 32)
    return out[:, 0] if like.ndim == 1 else out


def _brev_kv(m: np.ndarray, kv: int) -> np.ndarray:
    """Reversed kernel order: bit f of the result is bit kv - 1 - f of m.
    ``__brev(kmask) >> (32 - kv)`` for one word (cumm/conv/kernel.py:883-889);
    for dynamic masks the words are walked from the last with ``__brev`` and the
    bit offset ``used * 32 - kv`` (mma.py:201-265), the same map."""
    return _pack(mask_bits(m, kv)[:, ::-1], np.asarray(m))


def _filter_bits(mask_filter: int, kv: int) -> np.ndarray:
    """Offset f passes the filter when bit f % 32 of mask_filter is set
    (kernel.py:891-897 with ``filter_offset % 32`` for dynamic masks)."""
    return np.array([(mask_filter >> (f % 32)) & 1 for f in range(kv)], dtype=bool)


def conv_forward(X, W, pair, mask_sorted, argsort, tile_m: int, mask_filter: int = ALL,
                 alpha: float = 1.0, beta: float = 0.0, bias=None, out_init=None,
                 act_type: int = 0, act_alpha: float = 0.0, scale=None) -> Tuple[np.ndarray, np.ndarray]:
    """Forward (op 0): returns (output FP64 [N, K], mask_output [T] uint32).

    Tile t of sorted rows gets ``kmask_t = OR(mask[rows]) & mask_filter``
    (kernel.py:843-866), stored as mask_output[t] (868-874).  For each set bit f
    and row r of the tile, the gathered row ``X[pair[f][argsort[r]]]`` (zero for
    -1, sparse_iters.py:317-356, 629-631) is multiplied by ``W[:, f, :]^T``
    (mma.py:886-1005 / 445-578).  Every row is written, empty tiles included
    (kernel.py:907-920), at ``argsort[r]`` (out_iters.py:537-552) through the
    epilogue with source = bias (broadcast) or the output itself
    (conv/main.py:603-611, nvrtc_code.py:75-100).  ``scale`` (per output
    channel) multiplies the accumulator before alpha.
    """
    X = np.asarray(X, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    n = argsort.shape[0]
    K, kv, C = W.shape
    tm = tile_masks(mask_sorted, n, tile_m, mask_filter)
    row_bits = np.repeat(mask_bits(tm, kv), tile_m, axis=0)[:n]
    acc = np.zeros((n, K), dtype=np.float64)
    for f in range(kv):
        rows = np.nonzero(row_bits[:, f])[0]
        if rows.size == 0:
            continue
        idx = pair[f][argsort[rows]]
        ok = idx >= 0
        acc[rows[ok]] += X[idx[ok]] @ W[:, f, :].T
    if scale is not None:
        acc = acc * np.asarray(scale, dtype=np.float64)[None, :K]
    out = np.zeros((n, K)) if out_init is None else np.array(out_init, dtype=np.float64, copy=True)
    src = None
    if beta != 0.0:
        src = np.broadcast_to(np.asarray(bias, np.float64)[:K], (n, K)) if bias is not None else out[argsort]
    out[argsort] = epilogue(acc, alpha, beta, src, act_type, act_alpha)
    return out, tm.astype(np.uint32)


def conv_dgrad(dO, W, pair_bwd, mask_sorted, argsort, tile_m: int, mask_filter: int = ALL,
               reverse_mask: bool = False, alpha: float = 1.0, beta: float = 0.0, bias=None,
               din_init=None, act_type: int = 0, act_alpha: float = 0.0, scale=None) -> np.ndarray:
    """Dgrad (op 1): returns din FP64 [N, C].

    ``kmask_t = OR(mask[rows]) & mask_filter``; tiles with kmask_t == 0 return
    without writing (kernel.py:875-882); then, with ``reverse_mask``,
    ``kmask = brev(kmask) >> (32 - kv)`` (883-889).  Row r accumulates
    ``dO[pair_bwd[f][argsort[r]]] @ W[:, f, :]`` for each f in the tile mask and
    writes ``din[argsort[r]]``.
    """
    dO = np.asarray(dO, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    n = argsort.shape[0]
    K, kv, C = W.shape
    raw = tile_masks(mask_sorted, n, tile_m, mask_filter)
    live_t = mask_bits(raw, kv).any(axis=1)
    tm = _brev_kv(raw, kv) if reverse_mask else raw
    row_bits = np.repeat(mask_bits(tm, kv), tile_m, axis=0)[:n]
    row_live = np.repeat(live_t, tile_m)[:n]
    acc = np.zeros((n, C), dtype=np.float64)
    for f in range(kv):
        rows = np.nonzero(row_bits[:, f] & row_live)[0]
        if rows.size == 0:
            continue
        idx = pair_bwd[f][argsort[rows]]
        ok = idx >= 0
        acc[rows[ok]] += dO[idx[ok]] @ W[:, f, :]
    if scale is not None:
        acc = acc * np.asarray(scale, dtype=np.float64)[None, :C]
    din = np.zeros((n, C)) if din_init is None else np.array(din_init, dtype=np.float64, copy=True)
    live_rows = np.nonzero(row_live)[0]
    src = None
    if beta != 0.0:
        src = (np.broadcast_to(np.asarray(bias, np.float64)[:C], (live_rows.size, C)) if bias is not None
               else din[argsort[live_rows]])
    din[argsort[live_rows]] = epilogue(acc[live_rows], alpha, beta, src, act_type, act_alpha)
    return din


def conv_wgrad(X, dO, pair, mask_reduced, argsort, mask_width: int, mask_filter: int = ALL,
               alpha: float = 1.0, beta: float = 0.0, dW_init=None, kv: Optional[int] = None,
               act_type: int = 0, act_alpha: float = 0.0, bias=None) -> np.ndarray:
    """Wgrad (op 2): returns dW FP64 [K, kv, C].

    Offsets outside mask_filter are left untouched (kernel.py:891-897).  Sorted
    row r contributes to offset f only if offset f is set in
    ``mask_reduced[r // mask_width]`` (one word, or words [W] with dynamic masks) (the forward mask_output; the k-tile skip
    of mma.py:607-648 with mask_width % TK == 0) and ``pair[f][argsort[r]] != -1``:
    ``dW[:, f, :] = epi(sum_r dO[argsort[r]]^T X[pair[f][argsort[r]]])``, the
    split-K partitions summing into one result (kernel.py:960-1083).  A bias is
    a row-broadcast source indexed by the GEMM column, which for the
    ``[K, kv * C]`` view of the gradient is ``f * C + c`` (kernel.py:457-460).
    """
    X = np.asarray(X, dtype=np.float64)
    dO = np.asarray(dO, dtype=np.float64)
    n = argsort.shape[0]
    kv = pair.shape[0] if kv is None else kv
    K = dO.shape[1]
    C = X.shape[1]
    dW = np.zeros((K, kv, C)) if dW_init is None else np.array(dW_init, dtype=np.float64, copy=True)
    r = np.arange(n)
    mi = r // mask_width
    in_range = mi < mask_reduced.shape[0]
    gbits = mask_bits(mask_reduced, kv)
    rbits = np.zeros((n, kv), dtype=bool)
    rbits[in_range] = gbits[mi[in_range]]
    fbits = _filter_bits(mask_filter, kv)
    for f in range(kv):
        if not fbits[f]:
            continue
        rows = np.nonzero(rbits[:, f])[0]
        idx = pair[f][argsort[rows]]
        ok = idx >= 0
        acc = dO[argsort[rows[ok]]].T @ X[idx[ok]]
        src = None
        if beta != 0.0:
            src = (np.broadcast_to(np.asarray(bias, np.float64)[f * C:(f + 1) * C], (K, C)) if bias is not None
                   else dW[:, f, :])
        dW[:, f, :] = epilogue(acc, alpha, beta, src, act_type, act_alpha)
    return dW


# ---------------------------------------------------------------------------
# The reference's CPU path: compact pairs + per-offset gather / GEMM / scatter
# ---------------------------------------------------------------------------
def compact_pairs(pair: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """igemm table [kv, N_out] -> spconv's compact ``[2, kv, N]`` pairs and
    ``indice_num_per_loc`` (in = pairs[0], out = pairs[1]), the format the
    reference's numpy check consumes (cumm/conv/main_real_spconv.py:309-403)."""
    kv, n = pair.shape
    pairs = np.full((2, kv, n), -1, dtype=np.int32)
    nums = np.zeros(kv, dtype=np.int32)
    for f in range(kv):
        outs = np.nonzero(pair[f] >= 0)[0].astype(np.int32)
        nums[f] = outs.size
        pairs[0, f, :outs.size] = pair[f][outs]
        pairs[1, f, :outs.size] = outs
    return pairs, nums


def cpu_fwd(inp, W, pairs, nums, n_out: int) -> np.ndarray:
    """main_real_spconv.py:309-337 in the input dtype's FP32 arithmetic:
    ``out[o_f] += inp[i_f] @ W[:, f, :]^T`` per offset."""
    K, kv, C = W.shape
    weight_ref = np.ascontiguousarray(W.transpose(1, 0, 2))  # [kv, K, C]
    out = np.zeros((n_out, K), dtype=np.float32)
    for f in range(kv):
        nhot = nums[f]
        a_inds = pairs[0][f][:nhot]
        c_inds = pairs[1][f][:nhot]
        out[c_inds] += inp[a_inds].astype(np.float32) @ weight_ref[f].T.astype(np.float32)
    return out


def cpu_dgrad(dout, W, pairs, nums, n_in: int) -> np.ndarray:
    """main_real_spconv.py:347-372: ``din[i_f] += dout[o_f] @ W[:, f, :]``."""
    K, kv, C = W.shape
    weight_ref = np.ascontiguousarray(W.transpose(1, 0, 2))
    din = np.zeros((n_in, C), dtype=np.float32)
    for f in range(kv):
        nhot = nums[f]
        a_inds = pairs[1][f][:nhot]
        c_inds = pairs[0][f][:nhot]
        din[c_inds] += dout[a_inds].astype(np.float32) @ weight_ref[f].astype(np.float32)
    return din


def cpu_wgrad(inp, dout, pairs, nums, K: int, C: int) -> np.ndarray:
    """main_real_spconv.py:375-409: ``dW[:, f, :] = dout[o_f]^T @ inp[i_f]``."""
    kv = pairs.shape[1]
    dw = np.zeros((kv, K, C), dtype=np.float32)
    for f in range(kv):
        nhot = nums[f]
        o_inds = pairs[1][f][:nhot]
        i_inds = pairs[0][f][:nhot]
        dw[f] = dout[o_inds].astype(np.float32).T @ inp[i_inds].astype(np.float32)
    return np.ascontiguousarray(dw.transpose(1, 0, 2))


def relative_error(y: np.ndarray, y64: np.ndarray) -> Tuple[float, float]:
    """(max-abs, relative Frobenius) error of ``y`` against the FP64 reference."""
    y = np.asarray(y, dtype=np.float64)
    d = y - y64
    den = np.linalg.norm(y64)
    return float(np.max(np.abs(d))) if d.size else 0.0, float(np.linalg.norm(d) / den) if den > 0 else float(np.linalg.norm(d))
