"""Pins the CPU oracle (oracle/) and the host pair builder before anything is
compared against them.  Each check is chosen so a plausible mistake (dropped
term, wrong sign or index, transposed operand, wrong mask granularity) fails."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from oracle import bruteforce, gemm_ref, spconv_ref as ref
from cumm_amd import pointcloud as pc

ROOT = Path(__file__).resolve().parent.parent


def _tiny(n=220, grid=(6, 20, 20), seed=0, C=8, K=12):
    coords = pc.plane_scene(n, grid=grid, seed=seed)
    prob = pc.build_subm(coords, grid)
    rng = np.random.default_rng(seed + 7)
    X = rng.standard_normal((n, C))
    W = rng.standard_normal((K, 27, C))
    dO = rng.standard_normal((n, K))
    return prob, X, W, dO


# --------------------------------------------------------------- pair builder
def test_subm_pair_invariants():
    prob, *_ = _tiny()
    pf, n = prob.pair_fwd, prob.n
    assert np.array_equal(pf[13], np.arange(n)), "centre offset is the identity"
    for f in range(27):
        i = np.nonzero(pf[f] >= 0)[0]
        j = pf[f][i]
        # subm symmetry: pair[f][i] = j  <=>  pair[26 - f][j] = i
        assert np.array_equal(pf[26 - f][j], i)
    m = pc.row_masks(pf)
    for f in range(27):
        assert np.array_equal(((m >> f) & 1).astype(bool), pf[f] >= 0)
    assert np.array_equal(np.sort(prob.argsort), np.arange(n)), "argsort is a permutation"
    assert np.all(np.diff(prob.mask_sorted.astype(np.int64)) >= 0), "masks grouped by key"
    assert np.array_equal(prob.mask_sorted, prob.mask[prob.argsort])


def test_subm_pairs_match_bruteforce_neighbours():
    prob, *_ = _tiny(n=150, seed=3)
    lut = {tuple(c): i for i, c in enumerate(prob.coords.tolist())}
    offs = pc.kernel_offsets()
    for i, c in enumerate(prob.coords.tolist()):
        for f, o in enumerate(offs):
            nb = (c[0], c[1] + o[0] - 1, c[2] + o[1] - 1, c[3] + o[2] - 1)
            assert prob.pair_fwd[f][i] == lut.get(nb, -1)


def test_strided_pairs_inverse_and_bruteforce():
    coords = pc.plane_scene(200, grid=(6, 20, 20), seed=5)
    sp = pc.build_strided(coords, (6, 20, 20))
    for f in range(27):
        o = np.nonzero(sp.pair_fwd[f] >= 0)[0]
        assert np.array_equal(sp.pair_bwd[f][sp.pair_fwd[f][o]], o)
        assert np.count_nonzero(sp.pair_bwd[f] >= 0) == o.size
    # every output has at least one input; every input reaches an output
    assert np.all(pc.row_masks(sp.pair_fwd) != 0)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((200, 4))
    W = rng.standard_normal((5, 27, 4))
    out, _ = ref.conv_forward(X, W, sp.pair_fwd, sp.mask_fwd_sorted, sp.argsort_fwd, 16)
    bf = bruteforce.conv_forward_coords(coords, X, W, sp.out_coords, stride=2, padding=1, subm=False)
    np.testing.assert_allclose(out, bf, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------- forward
def test_forward_matches_bruteforce():
    prob, X, W, _ = _tiny()
    out, tm = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    bf = bruteforce.conv_forward_coords(prob.coords, X, W)
    np.testing.assert_allclose(out, bf, rtol=1e-12, atol=1e-12)
    assert np.array_equal(tm, pc.tile_masks(prob.mask_sorted, 16))


def test_forward_matches_reference_cpu_path():
    prob, X, W, _ = _tiny(seed=4)
    pairs, nums = ref.compact_pairs(prob.pair_fwd)
    out, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 32)
    cpu = ref.cpu_fwd(X.astype(np.float32), W.astype(np.float32), pairs, nums, prob.n)
    np.testing.assert_allclose(cpu, out, rtol=1e-4, atol=1e-4)


def test_forward_centre_only_is_dense_gemm():
    n, C, K = 100, 6, 5
    rng = np.random.default_rng(2)
    X, W = rng.standard_normal((n, C)), rng.standard_normal((K, 27, C))
    pair = np.full((27, n), -1, np.int32)
    pair[13] = np.arange(n)
    ms, order = pc.sort_masks(pc.row_masks(pair))
    out, _ = ref.conv_forward(X, W, pair, ms, order, 16)
    np.testing.assert_allclose(out, X @ W[:, 13, :].T, rtol=1e-12)
    # a row permutation of the pair column changes nothing
    perm = np.random.default_rng(3).permutation(n).astype(np.int32)
    out2, _ = ref.conv_forward(X, W, pair, ms, perm, 16)
    np.testing.assert_allclose(out2, out, rtol=1e-12)


def test_forward_all_missing_is_zero_and_written():
    n, C, K = 40, 4, 3
    pair = np.full((27, n), -1, np.int32)
    init = np.full((n, K), 7.0)
    out, tm = ref.conv_forward(np.ones((n, C)), np.ones((K, 27, C)), pair, np.zeros(n, np.uint32),
                               np.arange(n, dtype=np.int32), 16, out_init=init)
    assert np.all(out == 0.0) and np.all(tm == 0)
    out, _ = ref.conv_forward(np.ones((n, C)), np.ones((K, 27, C)), pair, np.zeros(n, np.uint32),
                              np.arange(n, dtype=np.int32), 16, out_init=init, beta=1.0)
    assert np.all(out == 7.0)


def test_forward_tile_mask_granularity():
    """A row whose own mask lacks f still gets f when a tile-mate has it
    (kmask is the OR over the tile, cumm/conv/kernel.py:843-866)."""
    n, C, K = 2, 3, 2
    pair = np.stack([np.arange(n, dtype=np.int32)] * 27)
    mask = np.array([1 << 0, 1 << 1], np.uint32)  # inconsistent on purpose
    order = np.arange(n, dtype=np.int32)
    rng = np.random.default_rng(0)
    X, W = rng.standard_normal((n, C)), rng.standard_normal((K, 27, C))
    out, tm = ref.conv_forward(X, W, pair, mask, order, 16)
    exp = X @ (W[:, 0, :] + W[:, 1, :]).T
    np.testing.assert_allclose(out, exp, rtol=1e-12)
    assert tm.tolist() == [3]
    out1, _ = ref.conv_forward(X, W, pair, mask, order, 1)  # tile of one row: own mask only
    np.testing.assert_allclose(out1, np.stack([X[0] @ W[:, 0, :].T, X[1] @ W[:, 1, :].T]), rtol=1e-12)
    outf, tmf = ref.conv_forward(X, W, pair, mask, order, 16, mask_filter=1)
    np.testing.assert_allclose(outf, X @ W[:, 0, :].T, rtol=1e-12)
    assert tmf.tolist() == [1]


def test_forward_epilogue():
    prob, X, W, _ = _tiny(seed=6)
    rng = np.random.default_rng(9)
    init = rng.standard_normal((prob.n, W.shape[0]))
    bias = rng.standard_normal(W.shape[0])
    plain, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    got, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, alpha=0.5, beta=2.0,
                              out_init=init)
    np.testing.assert_allclose(got, 0.5 * plain + 2.0 * init, rtol=1e-12)
    got, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, beta=1.0, bias=bias,
                              act_type=1)
    np.testing.assert_allclose(got, np.maximum(plain + bias, 0), rtol=1e-12)
    got, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, act_type=3,
                              act_alpha=0.1)
    np.testing.assert_allclose(got, np.where(plain >= 0, plain, 0.1 * plain), rtol=1e-12)
    got, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, act_type=2)
    np.testing.assert_allclose(got, 1 / (1 + np.exp(-plain)), rtol=1e-12)


def test_int8_inference_epilogue_known_answers():
    e = ref.int8_inference_epilogue
    one, zero = np.ones(1), np.zeros(1)
    col = lambda v: np.array(v, np.int64).reshape(-1, 1)
    # round half to even after scale * acc + bias
    got = e(col([0, 1, 2, 3, -1, -2, -3]), one, [0.5])
    assert got.ravel().tolist() == [0, 2, 2, 4, 0, -2, -2]
    # saturation at both ends, exact at the limits
    assert e(col([127, 128, 1000, -128, -129, -5000]), one, zero).ravel().tolist() == [127, 127, 127, -128, -128, -128]
    # scale then bias: 0.25 * 10 + 1.0 = 3.5 -> 4 ; 0.25 * 6 - 1 = 0.5 -> 0
    assert e(col([10, 6]), [0.25], [1.0]).ravel().tolist() == [4, 2]
    assert e(col([6]), [0.25], [-1.0]).ravel().tolist() == [0]
    # output_add with beta: 2 + 0.5 * 7 = 5.5 -> 6 ; alpha plays no part
    assert e(col([2]), one, zero, beta=0.5, source=col([7])).ravel().tolist() == [6]
    # activation before clamping: ReLU(-300) = 0 ; LeakyReLU(-300, 0.1) = -30
    assert e(col([-300]), one, zero, act_type=1).ravel().tolist() == [0]
    assert e(col([-300]), one, zero, act_type=3, act_alpha=0.1).ravel().tolist() == [-30]
    # per-column scale and bias are indexed by output channel
    acc = np.array([[4, 4, 4]], np.int64)
    assert e(acc, [1.0, 2.0, 0.5], [0.0, 1.0, -1.0]).ravel().tolist() == [4, 9, 1]


def test_int8_conv_equals_clipped_integer_bruteforce():
    prob, _, _, _ = _tiny(seed=12, C=16, K=8)
    rng = np.random.default_rng(3)
    X = rng.integers(-12, 13, (prob.n, 16))
    W = rng.integers(-6, 7, (8, 27, 16))
    acc, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    bf = bruteforce.conv_forward_coords(prob.coords, X, W)
    assert np.array_equal(acc, bf)
    got = ref.int8_inference_epilogue(acc.astype(np.int64), np.ones(8), np.zeros(8))
    assert np.array_equal(got, np.clip(bf, -128, 127).astype(np.int8))
    assert (np.abs(bf) > 127).any() and (np.abs(bf) < 127).any()


# --------------------------------------------------------------- dynamic (multi-word) masks
def _tiny5(n=160, seed=4, C=4, K=3, ks=5):
    grid = (6, 16, 16) if ks == 5 else (8, 16, 16)
    coords = pc.plane_scene(n, grid=grid, seed=seed)
    prob = pc.build_subm(coords, grid, ksize=(ks, ks, ks))
    rng = np.random.default_rng(seed + 3)
    return prob, rng.standard_normal((n, C)), rng.standard_normal((K, ks ** 3, C)), rng.standard_normal((n, K))


@pytest.mark.parametrize("ks", [5, 7])
def test_dynamic_masks_match_bruteforce(ks):
    """kv = 125 / 343: masks are [N, 4] / [N, 11] words; forward, dgrad
    (reversed masks) and wgrad (reduced multi-word masks) equal the
    coordinate brute force."""
    prob, X, W, dO = _tiny5(ks=ks, n=160 if ks == 5 else 120)
    kv, words = ks ** 3, (ks ** 3 + 31) // 32
    assert prob.kv == kv and prob.mask_sorted.shape == (prob.n, words)
    assert pc.mac_efficiency(prob.mask_sorted, 16, prob.num_pairs) <= 1.0
    out, tm = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    assert np.allclose(out, bruteforce.conv_forward_coords(prob.coords, X, W), rtol=1e-12, atol=1e-12)
    assert tm.shape == ((prob.n + 15) // 16, words) and np.array_equal(tm, pc.tile_masks(prob.mask_sorted, 16))
    din = ref.conv_dgrad(dO, W, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    assert np.allclose(din, bruteforce.conv_dgrad_coords(prob.coords, dO, W), rtol=1e-12, atol=1e-12)
    dw = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16)
    assert np.allclose(dw, bruteforce.conv_wgrad_coords(prob.coords, X, dO, kv), rtol=1e-12, atol=1e-12)


def test_dynamic_mask_filter_is_per_bit_position():
    """A filter bit b removes offsets b, b + 32, b + 64, b + 96 (filter_offset % 32)."""
    prob, X, W, dO = _tiny5(seed=9)
    filt = 0xFFFFFFFF & ~(1 << 5)
    out, tm = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, mask_filter=filt)
    W0 = W.copy()
    W0[:, [5, 37, 69, 101], :] = 0.0
    assert np.allclose(out, bruteforce.conv_forward_coords(prob.coords, X, W0), rtol=1e-12, atol=1e-12)
    assert not (tm & np.uint32(1 << 5)).any()
    dw = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16, mask_filter=filt)
    full = bruteforce.conv_wgrad_coords(prob.coords, X, dO, 125)
    keep = [f for f in range(125) if f % 32 != 5]
    assert np.allclose(dw[:, keep], full[:, keep]) and not dw[:, [5, 37, 69, 101]].any()


def test_multiword_representation_of_small_kernels_is_equivalent():
    """kv = 27 as [N, 1] words (a dynamic-mask layout) gives the 1-word results."""
    prob, X, W, dO = _tiny(seed=13)
    m2 = prob.mask_sorted.reshape(-1, 1)
    a, ta = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    b, tb = ref.conv_forward(X, W, prob.pair_fwd, m2, prob.argsort, 16)
    assert np.array_equal(a, b) and np.array_equal(ta, tb[:, 0])
    a = ref.conv_dgrad(dO, W, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    b = ref.conv_dgrad(dO, W, prob.pair_bwd, m2, prob.argsort, 16, reverse_mask=True)
    assert np.array_equal(a, b)


# --------------------------------------------------------------- dgrad / wgrad
def test_dgrad_adjoint_and_cpu_path():
    prob, X, W, dO = _tiny(seed=8)
    y, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    din = ref.conv_dgrad(dO, W, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    # <conv(X), dO> == <X, conv^T(dO)>
    assert np.isclose(np.sum(y * dO), np.sum(X * din), rtol=1e-11)
    mb, ab = pc.sort_masks(pc.row_masks(prob.pair_bwd))
    din2 = ref.conv_dgrad(dO, W, prob.pair_bwd, mb, ab, 16)
    np.testing.assert_allclose(din2, din, rtol=1e-12, atol=1e-12)
    pairs, nums = ref.compact_pairs(prob.pair_fwd)
    cpu = ref.cpu_dgrad(dO.astype(np.float32), W.astype(np.float32), pairs, nums, prob.n)
    np.testing.assert_allclose(cpu, din, rtol=1e-4, atol=1e-4)


def test_dgrad_skips_empty_tiles():
    n, C, K = 40, 3, 2
    pair = np.full((27, n), -1, np.int32)
    pair[13, :16] = np.arange(16)
    mask = pc.row_masks(pair)
    ms, order = pc.sort_masks(mask)  # the 24 empty rows sort first
    init = np.full((n, C), 5.0)
    din = ref.conv_dgrad(np.ones((n, K)), np.ones((K, 27, C)), pair, ms, order, 16, din_init=init)
    written = order[np.repeat(ref.tile_masks(ms, n, 16) != 0, 16)[:n]]
    untouched = np.setdiff1d(np.arange(n), written)
    assert untouched.size > 0 and np.all(din[untouched] == 5.0)


def test_wgrad_adjoint_and_cpu_path():
    prob, X, W, dO = _tiny(seed=10)
    tm = ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32)
    dW = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16)
    y, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    # d<conv(X; W), dO>/dW == wgrad, and the form is linear in W
    assert np.isclose(np.sum(y * dO), np.sum(W * dW), rtol=1e-11)
    pairs, nums = ref.compact_pairs(prob.pair_fwd)
    cpu = ref.cpu_wgrad(X.astype(np.float32), dO.astype(np.float32), pairs, nums, W.shape[0], X.shape[1])
    np.testing.assert_allclose(cpu, dW, rtol=1e-4, atol=1e-3)


def test_wgrad_reduced_mask_skip_and_filter():
    prob, X, W, dO = _tiny(seed=12)
    tm = ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32)
    full = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16)
    cut = tm.copy()
    cut[0] &= ~np.uint32(1 << 13)  # tile 0 no longer contributes to the centre offset
    part = ref.conv_wgrad(X, dO, prob.pair_fwd, cut, prob.argsort, 16)
    rows = prob.argsort[:16]
    delta = dO[rows].T @ X[prob.pair_fwd[13][rows]]
    np.testing.assert_allclose(full[:, 13, :] - part[:, 13, :], delta, rtol=1e-10, atol=1e-10)
    init = np.full_like(W, 3.0)
    filt = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16, mask_filter=0b111, dW_init=init)
    assert np.all(filt[:, 3:, :] == 3.0)
    np.testing.assert_allclose(filt[:, :3, :], full[:, :3, :])
    # coarser mask width: one mask word covers two 16-row tiles
    tm32 = ref.tile_masks(prob.mask_sorted, prob.n, 32).astype(np.uint32)
    np.testing.assert_allclose(ref.conv_wgrad(X, dO, prob.pair_fwd, tm32, prob.argsort, 32), full,
                               rtol=1e-12, atol=1e-12)


def test_wgrad_epilogue_bias_is_a_row_constant_source():
    """beta * bias in wgrad equals beta * (the output itself) when every row k
    of the output holds the same values bias[f * C + c]: the bias is indexed by
    the column of the [K, kv * C] view (a per-offset slice, not bias[c])."""
    prob, X, W, dO = _tiny(seed=14)
    K, kv, C = W.shape
    tm = ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(kv * C)
    init = np.broadcast_to(b.reshape(1, kv, C), (K, kv, C)).copy()
    via_bias = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16, alpha=0.5, beta=2.0,
                              dW_init=np.zeros_like(W), bias=b, act_type=3, act_alpha=0.1)
    via_out = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16, alpha=0.5, beta=2.0,
                             dW_init=init, act_type=3, act_alpha=0.1)
    np.testing.assert_allclose(via_bias, via_out, rtol=1e-13, atol=1e-13)
    # a bias read as bias[c] would repeat the first offset's slice: it must not
    wrong = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16, alpha=0.5, beta=2.0,
                           dW_init=np.broadcast_to(b[:C].reshape(1, 1, C), (K, kv, C)).copy(),
                           act_type=3, act_alpha=0.1)
    assert not np.allclose(via_bias, wrong)


# --------------------------------------------------------------- gemm
def test_gemm_closed_forms():
    rng = np.random.default_rng(12315)
    m, n, k = 70, 24, 36
    a = rng.uniform(-1, 1, (m, k))
    b = rng.uniform(-1, 1, (k, n))
    np.testing.assert_allclose(gemm_ref.matmul2(a, b, np.zeros((m, n))), a @ b)
    np.testing.assert_allclose(gemm_ref.matmul2(a.T.copy(), b.T.copy(), np.zeros((m, n)), trans_a=True,
                                                trans_b=True), a @ b)
    a_inds = rng.permutation(m).astype(np.int32)
    c_inds = rng.permutation(m).astype(np.int32)
    c = gemm_ref.matmul2(a, b, np.zeros((m, n)), shuffle=1, a_inds=a_inds, c_inds=c_inds)
    exp = np.zeros((m, n))
    exp[c_inds] = a[a_inds] @ b
    np.testing.assert_allclose(c, exp)
    b2 = rng.uniform(-1, 1, (m, n))
    c = gemm_ref.matmul2(a, b2, np.zeros((k, n)), trans_a=True, shuffle=2, a_inds=a_inds, b_inds=c_inds)
    np.testing.assert_allclose(c, a[a_inds].T @ b2[c_inds])
    # self-cancel known answer: A.B + 1 * (-A.B) == 0
    c = gemm_ref.matmul2(a, b, -(a @ b), beta=1.0)
    assert np.max(np.abs(c)) < 1e-12
    # bias broadcast
    bias = rng.uniform(-1, 1, n)
    np.testing.assert_allclose(gemm_ref.matmul2(a, b, np.zeros((m, n)), d=bias, beta=1.0), a @ b + bias)
    h = gemm_ref.hilbert(5, 5)
    assert np.isclose(h[2, 3], 1 / 6)


def test_py_matmul_blocked_fp32():
    rng = np.random.default_rng(19260817)
    a = rng.uniform(-1, 1, (33, 70)).astype(np.float32)
    b = rng.uniform(-1, 1, (70, 17)).astype(np.float32)
    np.testing.assert_allclose(gemm_ref.py_matmul(a, b), a.astype(np.float64) @ b, rtol=1e-5, atol=1e-5)


def _emu():
    lib_path = ROOT / "oracle" / "_build" / "libgemm_emu.so"
    if not lib_path.exists():
        pytest.skip("oracle/_build/libgemm_emu.so not built (make -C oracle)")
    L = ctypes.CDLL(str(lib_path))
    L.gemm_emu_f32.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 3
    return L


def test_gemm_emulator_config1():
    """Plain FP32 GEMM 128x128x64, U(-1,1): the CPU emulation of the FP32 kernel's
    accumulation order against numpy FP64 (the cudasim stand-in)."""
    L = _emu()
    rng = np.random.default_rng(12315)
    a = rng.uniform(-1, 1, (128, 64)).astype(np.float32)
    b = rng.uniform(-1, 1, (64, 128)).astype(np.float32)
    bt = np.ascontiguousarray(b.T)
    c = np.zeros((128, 128), np.float32)
    L.gemm_emu_f32(a.ctypes.data, bt.ctypes.data, c.ctypes.data, 128, 128, 64)
    ref64 = a.astype(np.float64) @ b.astype(np.float64)
    assert np.max(np.abs(c - ref64)) < 1e-4
    # a transposed operand or a dropped chunk would fail this:
    assert np.max(np.abs(c - (a.astype(np.float64) @ b.astype(np.float64)).T)) > 1e-2
    # ragged k (not a multiple of 16)
    k = 37
    a2, b2 = a[:, :k].copy(), np.ascontiguousarray(b[:k].T)
    c2 = np.zeros((128, 128), np.float32)
    L.gemm_emu_f32(a2.ctypes.data, b2.ctypes.data, c2.ctypes.data, 128, 128, k)
    assert np.max(np.abs(c2 - a2.astype(np.float64) @ b2.T.astype(np.float64))) < 1e-4


# --------------------------------------------------------------- golden fixtures
GOLD = ROOT / "tests" / "golden"


def test_golden_subm_fixture():
    """Pair builder bit-exact against the stored tables; the FP64 mask-walk
    oracle equal to the coordinate brute force stored in the fixture."""
    g = np.load(GOLD / "subm_tiny.npz")
    prob = pc.build_subm(g["coords"], (6, 14, 14))
    for name in ("pair_fwd", "pair_bwd", "mask", "mask_sorted", "argsort"):
        assert np.array_equal(getattr(prob, name), g[name]), name
    assert np.array_equal(pc.tile_masks(prob.mask_sorted, 16), g["tile_mask16"])
    X, W, dO = (g[k].astype(np.float64) for k in ("X", "W", "dO"))
    out, tm = ref.conv_forward(X, W, g["pair_fwd"], g["mask_sorted"], g["argsort"], 16)
    assert np.array_equal(tm, g["tile_mask16"])
    assert np.allclose(out, g["out"], rtol=0, atol=1e-12)
    din = ref.conv_dgrad(dO, W, g["pair_bwd"], g["mask_sorted"], g["argsort"], 16, reverse_mask=True)
    assert np.allclose(din, g["din"], rtol=0, atol=1e-12)
    dw = ref.conv_wgrad(X, dO, g["pair_fwd"], g["tile_mask16"], g["argsort"], 16)
    assert np.allclose(dw, g["dw"], rtol=0, atol=1e-11)


def test_golden_strided_fixture():
    g = np.load(GOLD / "strided_tiny.npz")
    sp = pc.build_strided(g["in_coords"], (6, 14, 14))
    for name in ("out_coords", "pair_fwd", "pair_bwd", "mask_fwd_sorted", "argsort_fwd", "mask_bwd_sorted",
                 "argsort_bwd"):
        assert np.array_equal(getattr(sp, name), g[name]), name
    X, W, dO = (g[k].astype(np.float64) for k in ("X", "W", "dO"))
    out, tm = ref.conv_forward(X, W, g["pair_fwd"], g["mask_fwd_sorted"], g["argsort_fwd"], 16)
    assert np.allclose(out, g["out"], rtol=0, atol=1e-12)
    din = ref.conv_dgrad(dO, W, g["pair_bwd"], g["mask_bwd_sorted"], g["argsort_bwd"], 16)
    assert np.allclose(din, g["din"], rtol=0, atol=1e-12)
    dw = ref.conv_wgrad(X, dO, g["pair_fwd"], tm, g["argsort_fwd"], 16)
    assert np.allclose(dw, g["dw"], rtol=0, atol=1e-11)


def test_golden_gemm_fixture():
    g = np.load(GOLD / "gemm_small.npz")
    got = gemm_ref.matmul2(g["a"], g["b"], np.zeros((128, 128)))
    assert np.allclose(got, g["c"], rtol=0, atol=1e-12)
    m, n = g["ac_c"].shape
    got = gemm_ref.matmul2(g["ac_a"], g["ac_b"], np.zeros((m, n)), shuffle=1, a_inds=g["ac_a_inds"],
                           c_inds=g["ac_c_inds"])
    assert np.allclose(got, g["ac_c"], rtol=0, atol=1e-12)
    got = gemm_ref.matmul2(g["ab_a"], g["ab_b"], np.zeros(g["ab_c"].shape), trans_a=True, shuffle=2,
                           a_inds=g["ab_a_inds"], b_inds=g["ab_b_inds"])
    assert np.allclose(got, g["ab_c"], rtol=0, atol=1e-12)


def _int_gemm_loops(A, B, alpha, beta, src, act, act_alpha, out_int8):
    """Brute force: Python integers, explicit loops, wrap and saturate spelled out."""
    def wrap(v):
        v &= 0xFFFFFFFF
        return v - (1 << 32) if v >= (1 << 31) else v
    ia, ib, iact = int(alpha), int(beta), int(act_alpha)  # int() truncates toward zero
    m, k = len(A), len(A[0])
    n = len(B[0])
    out = []
    for i in range(m):
        row = []
        for j in range(n):
            acc = 0
            for t in range(k):
                acc += int(A[i][t]) * int(B[t][j])
            x = wrap(ia * wrap(acc))
            if ib != 0:
                x = wrap(x + wrap(ib * int(src[i][j])))
            if act == 1 and x < 0:
                x = 0
            if act == 3 and x < 0:
                x = wrap(x * iact)
            if out_int8:
                x = -128 if x < -128 else (127 if x > 127 else x)
            row.append(x)
        out.append(row)
    return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("out_int8", [True, False])
@pytest.mark.parametrize("act", [0, 1, 3])
def test_int_gemm_oracle_brute_force(out_int8, act):
    rng = np.random.default_rng(41 + act)
    m, n, k = 5, 7, 48
    A = rng.integers(-128, 128, (m, k))
    B = rng.integers(-128, 128, (k, n))
    src = rng.integers(-128, 128, (m, n)).astype(np.int8 if out_int8 else np.int32)
    got = gemm_ref.matmul2_int(A, B, src, alpha=2.7, beta=-1.9, act_type=act, act_alpha=3.5)
    exp = _int_gemm_loops(A.tolist(), B.tolist(), 2.7, -1.9, src.tolist(), act, 3.5, out_int8)
    assert got.dtype == (np.int8 if out_int8 else np.int32)
    assert np.array_equal(got.astype(np.int64), exp)


def test_int_gemm_oracle_matches_float_form_without_saturation():
    """Small operands: no value saturates or wraps, so the int32 result equals
    the FP64 form exactly (alpha, beta integral)."""
    rng = np.random.default_rng(3)
    a = rng.integers(-3, 4, (40, 32))
    b = rng.integers(-3, 4, (32, 24))
    init = rng.integers(-5, 6, (40, 24)).astype(np.int32)
    got = gemm_ref.matmul2_int(a, b, init, alpha=1.0, beta=1.0)
    exp = gemm_ref.matmul2(a, b, init.astype(np.float64), alpha=1.0, beta=1.0)
    assert np.array_equal(got.astype(np.float64), exp)
    a_inds = rng.permutation(40)[:30]
    c_inds = rng.permutation(40)[:30]
    got = gemm_ref.matmul2_int(a, b, init.astype(np.int8), shuffle=1, a_inds=a_inds, c_inds=c_inds)
    exp = gemm_ref.matmul2(a, b, init.astype(np.float64), shuffle=1, a_inds=a_inds, c_inds=c_inds)
    assert np.array_equal(got.astype(np.float64), exp)
    got = gemm_ref.matmul2_int(a, a, np.zeros((32, 32), np.int32), trans_a=True, shuffle=2,
                               a_inds=a_inds, b_inds=c_inds)
    exp = a[a_inds].T.astype(np.float64) @ a[c_inds]
    assert np.array_equal(got.astype(np.float64), exp)


def test_int_gemm_oracle_known_answers():
    # saturation at both ends: 16 * 127 * 127 = 258064
    a = np.full((2, 16), 127)
    a[1] = -127
    b = np.full((16, 1), 127)
    assert gemm_ref.matmul2_int(a, b, np.zeros((2, 1), np.int8)).ravel().tolist() == [127, -128]
    assert gemm_ref.matmul2_int(a, b, np.zeros((2, 1), np.int32)).ravel().tolist() == [258064, -258064]
    # alpha and beta truncate toward zero: 1.9 -> 1, -0.5 -> 0 (the source is not read)
    assert gemm_ref.matmul2_int(a, b, np.zeros((2, 1), np.int32), alpha=1.9).ravel().tolist() == [258064, -258064]
    assert gemm_ref.matmul2_int(a, b, np.full((2, 1), 7, np.int32), alpha=-0.5, beta=0.5).ravel().tolist() == [0, 0]
    # int32 wrap: 2^30 * 4 = 2^32 -> 0; LeakyReLU with act_alpha 2.9 -> 2
    one = np.ones((1, 16))
    zero = np.zeros((16, 1))
    assert gemm_ref.matmul2_int(one, zero, np.full((1, 1), 4, np.int32), beta=float(1 << 30)).item() == 0
    assert gemm_ref.matmul2_int(a[1:], b, np.zeros((1, 1), np.int32), act_type=3, act_alpha=2.9).item() == -516128
    with pytest.raises(ValueError):
        gemm_ref.matmul2_int(a, b, np.zeros((2, 1), np.int8), act_type=2)
